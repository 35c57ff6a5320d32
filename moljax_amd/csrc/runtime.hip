// Library runtime: error reporting and the per-device FFT twiddle tables.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>

#include "mjx_common.h"
#include "mjx_fft.h"

namespace mjx {

// Entry i of the tables: exp(-2*pi*i*t/n) for the power-of-two table
// (n = MJX_TWN, t = i) or the mixed-radix length whose table holds i
// (t = i - tw_base(n)), evaluated in FP64 (sincospi) and rounded once to FP32.
__global__ void k_init_twiddles(cpx<double>* tw64, cpx<float>* tw32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MJX_TW_TOTAL) return;
  int n = MJX_TWN, t = i;
#define X(N)                                  \
  if (i >= tw_base(N) && i < tw_base(N) + N) { \
    n = N;                                    \
    t = i - tw_base(N);                       \
  }
  MJX_FOR_MIXED(X)
#undef X
  double s, c;
  sincospi(2.0 * (double)t / (double)n, &s, &c);
  tw64[i] = mk(c, -s);
  tw32[i] = mk((float)c, (float)-s);
}

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MJX_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return MJX_OK;
}

static constexpr int kMaxDev = 64;
static std::mutex g_init_mu;
static cpx<double>* g_tw64[kMaxDev] = {nullptr};
static cpx<float>* g_tw32[kMaxDev] = {nullptr};

static int current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return -1;
  return dev;
}

// Allocates and fills the current device's tables on first use (a one-time,
// synchronous set-up; later calls only read the cached pointers).
int ensure_tables(hipStream_t) {
  const int dev = current_device();
  if (dev < 0) return fail(MJX_E_HIP, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_tw64[dev]) return MJX_OK;
  cpx<double>* t64 = nullptr;
  cpx<float>* t32 = nullptr;
  if (hipMalloc((void**)&t64, sizeof(cpx<double>) * MJX_TW_TOTAL) != hipSuccess)
    return fail(MJX_E_HIP, "hipMalloc twiddles");
  if (hipMalloc((void**)&t32, sizeof(cpx<float>) * MJX_TW_TOTAL) != hipSuccess) {
    (void)hipFree(t64);
    return fail(MJX_E_HIP, "hipMalloc twiddles");
  }
  hipLaunchKernelGGL(k_init_twiddles, dim3((MJX_TW_TOTAL + 255) / 256), dim3(256), 0, 0, t64, t32);
  int rc = check_launch("init_twiddles");
  if (rc) return rc;
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(MJX_E_HIP, std::string("init_twiddles sync: ") + hipGetErrorString(e));
  g_tw32[dev] = t32;
  g_tw64[dev] = t64;
  return MJX_OK;
}

bool tables_ready() {
  const int dev = current_device();
  if (dev < 0) return false;
  std::lock_guard<std::mutex> lk(g_init_mu);
  return g_tw64[dev] != nullptr && g_tw32[dev] != nullptr;
}

const cpx<float>* twiddles_f32() {
  const int dev = current_device();
  return dev < 0 ? nullptr : g_tw32[dev];
}
const cpx<double>* twiddles_f64() {
  const int dev = current_device();
  return dev < 0 ? nullptr : g_tw64[dev];
}

}  // namespace mjx

namespace mjx {

// [kind][dtype][log2 n], lock-free: a launch reads one entry.
static std::atomic<int> g_tuned[TK_NUM][2][13];

static int log2_len(int n) {
  int k = 0;
  while ((1 << k) < n && k < 12) ++k;
  return (1 << k) == n ? k : -1;
}

int tuned_lines(int kind, int dtype, int n) {
  const int k = log2_len(n);
  if (kind < 0 || kind >= TK_NUM || (dtype != MJX_F32 && dtype != MJX_F64) || k < 0) return 0;
  return g_tuned[kind][dtype][k].load(std::memory_order_relaxed);
}

void set_tuned_lines(int kind, int dtype, int n, int lines) {
  const int k = log2_len(n);
  if (kind < 0 || kind >= TK_NUM || (dtype != MJX_F32 && dtype != MJX_F64) || k < 0) return;
  g_tuned[kind][dtype][k].store(lines, std::memory_order_relaxed);
}

}  // namespace mjx

extern "C" {

int mjx_tune_get(int kind, int dtype, int n) { return mjx::tuned_lines(kind, dtype, n); }

int mjx_tune_set(int kind, int dtype, int n, int lines) {
  if (kind < 0 || kind >= mjx::TK_NUM) return mjx::fail(MJX_E_ARG, "unknown tuning kind");
  if (dtype != MJX_F32 && dtype != MJX_F64) return mjx::fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (!mjx::is_pow2_len(n) && n != 8) return mjx::fail(MJX_E_ARG, "n must be a power of two in [16, 4096]");
  if (lines < 0 || lines > 64 || (lines & (lines - 1))) return mjx::fail(MJX_E_ARG, "lines must be 0 or a power of two <= 64");
  mjx::set_tuned_lines(kind, dtype, n, lines);
  return MJX_OK;
}

int mjx_version(void) { return 100; }

const char* mjx_last_error(void) { return mjx::g_err.c_str(); }

int mjx_init(void) {
  const int rc = mjx::ensure_tables(nullptr);
  return rc ? rc : mjx::scratch_setup();
}

}  // extern "C"
