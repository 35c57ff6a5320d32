// C-ABI entry points of the periodic spectral path (kernels in
// spectral_kernels.h, instantiated in spectral_f32.hip / spectral_f64.hip).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "mjx_common.h"
#include "czt_api.h"
#include "gs_chain.h"
#include "spectral_api.h"

namespace mjx {

// Opt-in per-kernel timing: when enabled, every launch of the two hot kernels
// of the fused IMEX chain (row_gs, col_apply) is bracketed by HIP events on
// its own stream; totals are read back by mjx_kernel_timing().
enum { KT_ROW_GS = 0, KT_COL_APPLY = 1, KT_NUM = 2 };
struct Timed {
  int kind;
  hipEvent_t e0, e1;
};
static std::mutex g_kt_mu;
static bool g_kt_on = false;
static std::vector<Timed> g_kt_pending;
static double g_kt_ms[KT_NUM] = {0, 0};
static long long g_kt_n[KT_NUM] = {0, 0};
static std::vector<double> g_kt_samples[KT_NUM];

// Resolve the pending event pairs into totals and per-launch samples.
static int kt_resolve() {
  for (auto& t : g_kt_pending) {
    if (hipEventSynchronize(t.e1) != hipSuccess) return fail(MJX_E_HIP, "hipEventSynchronize failed");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.e0, t.e1) != hipSuccess) return fail(MJX_E_HIP, "hipEventElapsedTime failed");
    g_kt_ms[t.kind] += ms;
    g_kt_n[t.kind] += 1;
    g_kt_samples[t.kind].push_back(ms);
    (void)hipEventDestroy(t.e0);
    (void)hipEventDestroy(t.e1);
  }
  g_kt_pending.clear();
  return MJX_OK;
}

template <class Fn>
static int timed_launch(int kind, hipStream_t st, Fn&& fn) {
  bool on;
  {
    std::lock_guard<std::mutex> lk(g_kt_mu);
    on = g_kt_on;
  }
  if (!on) return fn();
  Timed t;
  t.kind = kind;
  if (hipEventCreate(&t.e0) != hipSuccess || hipEventCreate(&t.e1) != hipSuccess)
    return fail(MJX_E_HIP, "hipEventCreate failed");
  (void)hipEventRecord(t.e0, st);
  const int rc = fn();
  (void)hipEventRecord(t.e1, st);
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt_pending.push_back(t);
  return rc;
}

static int check_grid(int dtype, int B, int ny, int nx) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (B < 1) return fail(MJX_E_ARG, "batch must be >= 1");
  if (spec_generic(nx, ny) && (!czt_len_ok(nx) || !czt_len_ok(ny)))
    return fail(MJX_E_SIZE, "grid " + std::to_string(ny) + "x" + std::to_string(nx) +
                                " not supported: nx and ny must be in [2, 2048], or both Stockham lengths "
                                "(powers of two in [16, 4096], 96, 100, 192, 200, 250, 384, 500, 768, 1000)");
  if ((long long)B * ny * (nx / 16) > 0x7fffffffLL) return fail(MJX_E_LIMIT, "batch too large for one launch");
  return MJX_OK;
}

template <typename T>
static int spectral_pair_t(int B, int ny, int nx, const T* a_in, const T* b_in, long long ip_, long long is_,
                           T* a_out, T* b_out, long long op_, long long os_, int ghosts, const T* lamx,
                           const T* lamy, const T* lam2d, const InstParams* inst, int nrep, cpx<T>* W,
                           hipStream_t st) {
  int rc = spec_row_fwd_pack(nx, B, ny, a_in, b_in, ip_, is_, W, st);
  if (rc) return rc;
  rc = spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, nrep, st);
  if (rc) return rc;
  return spec_row_inv_unpack(nx, B, ny, W, a_out, b_in ? b_out : nullptr, op_, os_, ghosts, st);
}

// timed: the opt-in per-kernel events may bracket the hot launches (the
// allocation-free entry point passes false: it creates no events).
template <typename T>
static int gs_imex_t(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                     const T* lamx, const T* lamy, const T* lam2d, const InstParams* inst, int n_steps, cpx<T>* W,
                     hipStream_t st, bool timed = true) {
  int rc = spec_row_fwd_pack(nx, B, ny, u, v, pitch, istride, W, st);
  if (rc) return rc;
  rc = spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 1, st);
  if (rc) return rc;
  for (int s = 0; s < n_steps; ++s) {
    auto row = [&] { return spec_row_gs(nx, B, ny, W, inst, st); };
    rc = timed ? timed_launch(KT_ROW_GS, st, row) : row();
    if (rc) return rc;
    const int nrep = (s == n_steps - 1) ? 1 : 2;
    auto col = [&] { return spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, nrep, st); };
    rc = timed ? timed_launch(KT_COL_APPLY, st, col) : col();
    if (rc) return rc;
  }
  return spec_row_inv_unpack(nx, B, ny, W, u, v, pitch, istride, ghosts, st);
}

// The same operator in increment form (GS_STRANG_INC, gs_chain.h), in place;
// the step's row pass and column pass take the per-kernel events.
template <typename T>
static int gs_imex_inc_t(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                         const T* lamx, const T* lamy, const T* lam2d, const InstParams* inst, int n_steps,
                         cpx<T>* W, hipStream_t st) {
  int rc = spec_row_gs_inc(nx, B, ny, W, u, v, pitch, istride, inst, INC_FIRST, 0, st);
  if (!rc) rc = spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, NREP_INC1, st);
  for (int s = 0; s < n_steps && !rc; ++s) {
    rc = timed_launch(KT_ROW_GS, st,
                      [&] { return spec_row_gs_inc(nx, B, ny, W, u, v, pitch, istride, inst, INC_MID, 0, st); });
    const int nrep = s == n_steps - 1 ? NREP_INC1 : NREP_INC2;
    if (!rc)
      rc = timed_launch(KT_COL_APPLY, st,
                        [&] { return spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, nrep, st); });
  }
  if (rc) return rc;
  return spec_row_gs_inc(nx, B, ny, W, u, v, pitch, istride, inst, INC_LAST, ghosts, st);
}

// Strang chain whose instances take their own number of steps (instance row
// aux, 0..max_steps): the masked column and row passes leave an instance
// untouched once its steps are done, so each instance's arithmetic is that of
// a chain of its own length.
template <typename T>
static int gs_imex_var_t(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                         const T* lamx, const T* lamy, const T* lam2d, const InstParams* inst, int max_steps,
                         cpx<T>* W, hipStream_t st) {
  int rc = spec_row_fwd_pack(nx, B, ny, u, v, pitch, istride, W, st);
  if (!rc) rc = spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 1, st);
  for (int s = 0; s < max_steps && !rc; ++s) {
    rc = spec_row_gs(nx, B, ny, W, inst, st, RXN_HEUN, s);
    if (!rc) rc = spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 0, st, s);
  }
  if (rc) return rc;
  return spec_row_inv_unpack(nx, B, ny, W, u, v, pitch, istride, ghosts, st);
}

// ---------------------------------------------- allocation-free entry point
// The spectral workspace of mjx_gs_imex_strang_noalloc lives in this
// file-scope device array: one copy per GPU, held until the process ends.
// In host code `scratch` only names the symbol; its device address on the
// current GPU comes from hipGetSymbolAddress (scratch_setup) and that
// address, never the array's name, is what the kernels receive.
static __device__ __align__(256) unsigned char scratch[MJX_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {nullptr};

// Resolves the current device's address of `scratch` once per device (called
// by mjx_init, and by the entry point's first call on a device).
int scratch_setup() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail(MJX_E_HIP, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return MJX_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return fail(MJX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  return MJX_OK;
}

static void* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_spectral_workspace_bytes(int dtype, int B, int ny, int nx) {
  const long long es = dtype == MJX_F64 ? 16 : 8;
  return (long long)B * ny * spec_work_cols(nx, ny) * es;
}

int mjx_spectral_pair(int dtype, int B, int ny, int nx, const void* a_in, const void* b_in, long long in_pitch,
                      long long in_istride, void* a_out, void* b_out, long long out_pitch, long long out_istride,
                      int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d, const double* inst,
                      int nrep, void* work, void* stream) {
  int rc = check_grid(dtype, B, ny, nx);
  if (rc) return rc;
  if (!a_in || !a_out || !work || !inst) return fail(MJX_E_ARG, "null pointer argument");
  if (!lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (b_in && !b_out) return fail(MJX_E_ARG, "b_out required when b_in is given");
  if (nrep < NREP_PER_INSTANCE) return fail(MJX_E_ARG, "nrep must be >= -3");
  if (in_pitch < nx || out_pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && out_pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs out_pitch >= nx + 2");
  rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (dtype == MJX_F32)
    return spectral_pair_t<float>(B, ny, nx, (const float*)a_in, (const float*)b_in, in_pitch, in_istride,
                                  (float*)a_out, (float*)b_out, out_pitch, out_istride, fill_ghosts,
                                  (const float*)lamx, (const float*)lamy, (const float*)lam2d, ip, nrep,
                                  (cpx<float>*)work, st);
  return spectral_pair_t<double>(B, ny, nx, (const double*)a_in, (const double*)b_in, in_pitch, in_istride,
                                 (double*)a_out, (double*)b_out, out_pitch, out_istride, fill_ghosts,
                                 (const double*)lamx, (const double*)lamy, (const double*)lam2d, ip, nrep,
                                 (cpx<double>*)work, st);
}

int mjx_gs_imex_strang(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                       int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d, const double* inst,
                       int n_steps, void* work, void* stream) {
  int rc = check_grid(dtype, B, ny, nx);
  if (rc) return rc;
  if (!u || !v || !inst || !work) return fail(MJX_E_ARG, "null pointer argument");
  if (!lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  if (n_steps == 0) return MJX_OK;
  rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (dtype == MJX_F32)
    return gs_imex_t<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)lamx,
                            (const float*)lamy, (const float*)lam2d, ip, n_steps, (cpx<float>*)work, st);
  return gs_imex_t<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)lamx,
                           (const double*)lamy, (const double*)lam2d, ip, n_steps, (cpx<double>*)work, st);
}

int mjx_gs_imex_strang_noalloc(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch,
                               long long istride, int fill_ghosts, const void* lamx, const void* lamy,
                               const void* lam2d, const double* inst, int n_steps, void* stream) {
  int rc = check_grid(dtype, B, ny, nx);
  if (rc) return rc;
  if (spec_generic(nx, ny))
    return fail(MJX_E_SIZE,
                "mjx_gs_imex_strang_noalloc serves Stockham sides only (powers of two in [16, 4096] and the "
                "mixed-radix lengths 96, 100, 192, 200, 250, 384, 500, 768, 1000)");
  if (!u || !v || !inst) return fail(MJX_E_ARG, "null pointer argument");
  if (!lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  const long long need = mjx_spectral_workspace_bytes(dtype, B, ny, nx);
  if (need > MJX_NOALLOC_SCRATCH_BYTES)
    return fail(MJX_E_LIMIT, "mjx_gs_imex_strang_noalloc: the workspace needs " + std::to_string(need) +
                                 " bytes, above the fixed scratch of " + std::to_string(MJX_NOALLOC_SCRATCH_BYTES));
  if (n_steps == 0) return MJX_OK;
  if (!tables_ready())
    return fail(MJX_E_ARG, "mjx_gs_imex_strang_noalloc: call mjx_init on this device first");
  if ((rc = scratch_setup())) return rc;
  void* W = scratch_ptr();
  if (!W) return fail(MJX_E_HIP, "scratch address unavailable");
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  // every workspace element is written by the first pass before any read,
  // so the scratch needs no clearing between calls
  if (dtype == MJX_F32)
    return gs_imex_t<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)lamx,
                            (const float*)lamy, (const float*)lam2d, ip, n_steps, (cpx<float>*)W, st, false);
  return gs_imex_t<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)lamx,
                           (const double*)lamy, (const double*)lam2d, ip, n_steps, (cpx<double>*)W, st, false);
}

int mjx_gs_imex_strang_var(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                           int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d,
                           const double* inst, int max_steps, void* work, void* stream) {
  int rc = check_grid(dtype, B, ny, nx);
  if (rc) return rc;
  if (!u || !v || !inst || !work) return fail(MJX_E_ARG, "null pointer argument");
  if (!lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (max_steps < 0) return fail(MJX_E_ARG, "max_steps must be >= 0");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  if (max_steps == 0) return MJX_OK;
  rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (dtype == MJX_F32)
    return gs_imex_var_t<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)lamx,
                                (const float*)lamy, (const float*)lam2d, ip, max_steps, (cpx<float>*)work, st);
  return gs_imex_var_t<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)lamx,
                               (const double*)lamy, (const double*)lam2d, ip, max_steps, (cpx<double>*)work, st);
}

int mjx_gs_imex(int dtype, int scheme, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d, const double* inst,
                int n_steps, void* work, void* stream) {
  if (scheme == GS_STRANG)
    return mjx_gs_imex_strang(dtype, B, ny, nx, u, v, pitch, istride, fill_ghosts, lamx, lamy, lam2d, inst,
                              n_steps, work, stream);
  if (scheme != GS_EULER && scheme != GS_SSPRK2 && scheme != GS_BENCH_STRANG && scheme != GS_STRANG_INC)
    return fail(MJX_E_ARG,
                "scheme must be 0 (Strang), 1 (Euler), 2 (SSPRK2), 3 (benchmark Strang) or 4 (Strang, increment form)");
  int rc = check_grid(dtype, B, ny, nx);
  if (rc) return rc;
  if (scheme == GS_STRANG_INC && (!is_pow2_len(nx) || !is_pow2_len(ny)))
    return fail(MJX_E_SIZE, "increment-form IMEX-Strang: nx and ny must be powers of two in [16, 4096]");
  if (!u || !v || !inst || !work) return fail(MJX_E_ARG, "null pointer argument");
  if (!lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  if (n_steps == 0) return MJX_OK;
  rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (scheme == GS_STRANG_INC) {
    if (dtype == MJX_F32)
      return gs_imex_inc_t<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)lamx,
                                  (const float*)lamy, (const float*)lam2d, ip, n_steps, (cpx<float>*)work, st);
    return gs_imex_inc_t<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts,
                                 (const double*)lamx, (const double*)lamy, (const double*)lam2d, ip, n_steps,
                                 (cpx<double>*)work, st);
  }
  if (dtype == MJX_F32)
    return gs_chain_t<float>(scheme, B, ny, nx, (float*)u, (float*)v, pitch, istride, (float*)u, (float*)v, pitch,
                             istride, fill_ghosts, (const float*)lamx, (const float*)lamy, (const float*)lam2d, ip,
                             n_steps, (cpx<float>*)work, st);
  return gs_chain_t<double>(scheme, B, ny, nx, (double*)u, (double*)v, pitch, istride, (double*)u, (double*)v, pitch,
                            istride, fill_ghosts, (const double*)lamx, (const double*)lamy, (const double*)lam2d, ip,
                            n_steps, (cpx<double>*)work, st);
}

int mjx_kernel_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt_on = on != 0;
  if (on) {
    for (auto& t : g_kt_pending) {
      (void)hipEventDestroy(t.e0);
      (void)hipEventDestroy(t.e1);
    }
    g_kt_pending.clear();
    for (int i = 0; i < KT_NUM; ++i) {
      g_kt_ms[i] = 0;
      g_kt_n[i] = 0;
      g_kt_samples[i].clear();
    }
  }
  return MJX_OK;
}

int mjx_kernel_timing(int kind, double* total_ms, long long* count) {
  if (kind < 0 || kind >= KT_NUM || !total_ms || !count) return fail(MJX_E_ARG, "bad kernel timing query");
  std::lock_guard<std::mutex> lk(g_kt_mu);
  const int rc = kt_resolve();
  if (rc) return rc;
  *total_ms = g_kt_ms[kind];
  *count = g_kt_n[kind];
  return MJX_OK;
}

long long mjx_kernel_timing_samples(int kind, double* out, long long cap) {
  if (kind < 0 || kind >= KT_NUM) return fail(MJX_E_ARG, "bad kernel timing query");
  std::lock_guard<std::mutex> lk(g_kt_mu);
  const int rc = kt_resolve();
  if (rc) return rc;
  const auto& v = g_kt_samples[kind];
  const long long n = (long long)v.size();
  for (long long i = 0; out && i < n && i < cap; ++i) out[i] = v[(size_t)i];
  return n;
}

}  // extern "C"
