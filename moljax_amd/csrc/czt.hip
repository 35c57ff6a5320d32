// Host side of the any-length (chirp-z) path: plan tables and launchers for
// the periodic spectral passes and the non-periodic line transforms.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "czt_api.h"
#include "czt_kernels.h"

namespace mjx {

namespace {

// One plan: n points, generalised DFT of period P with doubled offsets
// K2 = 2 ko (output) and N2 = 2 no (input).  Tables [c (n) | d (n) | e (L)].
struct HostPlan {
  int n = 0, L = 0, lg = 0;
  cpx<float>* t32 = nullptr;
  cpx<double>* t64 = nullptr;
};

std::mutex g_mu;
std::map<std::array<int, 5>, HostPlan> g_plans;

struct C64 {
  double re, im;
};

// exp(-i pi a^2 / (4P)) with a^2 reduced exactly modulo 8P.
C64 chirp(long long a, int P) {
  const long long m = 8LL * P;
  long long r = a % m;
  if (r < 0) r += m;
  const long long q = (r * r) % m;
  const double th = M_PI * (double)q / (4.0 * (double)P);
  return {std::cos(th), -std::sin(th)};
}

// In-place forward FFT (exp(-2 pi i jk / L)) of a power-of-two length.
void host_fft(std::vector<C64>& x) {
  const int L = (int)x.size();
  for (int i = 1, j = 0; i < L; ++i) {
    int bit = L >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
  for (int len = 2; len <= L; len <<= 1) {
    for (int i = 0; i < L; i += len) {
      for (int k = 0; k < len / 2; ++k) {
        const double th = -2.0 * M_PI * (double)k / (double)len;
        const C64 w{std::cos(th), std::sin(th)};
        const C64 u = x[i + k], v = x[i + k + len / 2];
        const C64 t{w.re * v.re - w.im * v.im, w.re * v.im + w.im * v.re};
        x[i + k] = {u.re + t.re, u.im + t.im};
        x[i + k + len / 2] = {u.re - t.re, u.im - t.im};
      }
    }
  }
}

int get_plan(int n, int P, int K2, int N2, HostPlan& out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(MJX_E_HIP, "hipGetDevice failed");
  const std::array<int, 5> key{dev, n, P, K2, N2};
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    out = it->second;
    return MJX_OK;
  }
  HostPlan h;
  h.n = n;
  h.lg = 1;
  while ((1 << h.lg) < 2 * n - 1) ++h.lg;
  h.L = 1 << h.lg;
  if (h.L > MJX_TWN) return fail(MJX_E_SIZE, "chirp-z length " + std::to_string(n) + " too large");
  std::vector<C64> tab(2 * (size_t)n + h.L, C64{0.0, 0.0});
  for (int k = 0; k < n; ++k) tab[k] = chirp(2LL * k + K2, P);
  for (int i = 0; i < n; ++i) tab[n + i] = chirp(2LL * i + N2, P);
  std::vector<C64> e(h.L, C64{0.0, 0.0});
  for (int m = -(n - 1); m <= n - 1; ++m) {
    const C64 w = chirp(2LL * m + K2 - N2, P);
    e[(m + h.L) % h.L] = {w.re, -w.im};
  }
  host_fft(e);
  for (int k = 0; k < h.L; ++k) tab[2 * (size_t)n + k] = {e[k].re / h.L, e[k].im / h.L};
  const size_t cnt = tab.size();
  std::vector<cpx<float>> t32(cnt);
  std::vector<cpx<double>> t64(cnt);
  for (size_t i = 0; i < cnt; ++i) {
    t64[i].re = tab[i].re;
    t64[i].im = tab[i].im;
    t32[i].re = (float)tab[i].re;
    t32[i].im = (float)tab[i].im;
  }
  if (hipMalloc(&h.t32, cnt * sizeof(cpx<float>)) != hipSuccess ||
      hipMalloc(&h.t64, cnt * sizeof(cpx<double>)) != hipSuccess)
    return fail(MJX_E_HIP, "chirp-z table allocation failed");
  if (hipMemcpy(h.t32, t32.data(), cnt * sizeof(cpx<float>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h.t64, t64.data(), cnt * sizeof(cpx<double>), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MJX_E_HIP, "chirp-z table upload failed");
  g_plans[key] = h;
  out = h;
  return MJX_OK;
}

template <typename T>
CztPlan<T> dev_plan(const HostPlan& h) {
  CztPlan<T> p;
  const cpx<T>* t;
  if constexpr (std::is_same<T, float>::value) t = h.t32;
  else t = h.t64;
  p.c = t;
  p.d = t + h.n;
  p.e = t + 2 * h.n;
  p.n = h.n;
  p.L = h.L;
  p.lg = h.lg;
  return p;
}

// Threads per workgroup: one per butterfly up to 256, at least one wave.
int block_threads(int L) {
  const int h = L / 2;
  return h < 64 ? 64 : (h > 256 ? 256 : h);
}

template <typename T>
size_t smem_bytes(int L) {
  return (size_t)L * sizeof(cpx<T>);
}

int grid_check(long long blocks) {
  if (blocks < 1 || blocks > 0x7fffffffLL) return fail(MJX_E_LIMIT, "chirp-z launch too large");
  return MJX_OK;
}

}  // namespace

// ---------------------------------------------------------------- periodic
template <typename T>
static int row_fwd_pack_t(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,
                          cpx<T>* W, hipStream_t st, double in_scale) {
  HostPlan h;
  int rc = get_plan(nx, nx, 0, 0, h);
  if (rc || (rc = grid_check((long long)B * ny))) return rc;
  hipLaunchKernelGGL(k_czt_row_fwd_pack<T>, dim3(B * ny), dim3(block_threads(h.L)), smem_bytes<T>(h.L), st, a, b,
                     pitch, istride, W, ny, nx / 2 + 1, (T)in_scale, dev_plan<T>(h), twiddles<T>());
  return check_launch("czt_row_fwd_pack");
}

template <typename T>
static int row_inv_unpack_t(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,
                            long long istride, int ghosts, hipStream_t st, double keep, const InstParams* react) {
  HostPlan h;
  int rc = get_plan(nx, nx, 0, 0, h);
  if (rc || (rc = grid_check((long long)B * ny))) return rc;
  hipLaunchKernelGGL(k_czt_row_inv_unpack<T>, dim3(B * ny), dim3(block_threads(h.L)), smem_bytes<T>(h.L), st, W, a,
                     b, pitch, istride, ny, nx / 2 + 1, ghosts, (T)keep, react, dev_plan<T>(h), twiddles<T>());
  return check_launch("czt_row_inv_unpack");
}

template <typename T>
static int row_gs_t(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) {
  HostPlan h;
  int rc = get_plan(nx, nx, 0, 0, h);
  if (rc || (rc = grid_check((long long)B * ny))) return rc;
  hipLaunchKernelGGL(k_czt_row_gs<T>, dim3(B * ny), dim3(block_threads(h.L)), smem_bytes<T>(h.L), st, W, ip, ny,
                     nx / 2 + 1, rxn, mstep, dev_plan<T>(h), twiddles<T>());
  return check_launch("czt_row_gs");
}

template <typename T>
static int col_apply_t(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                       const InstParams* ip, int nrep, hipStream_t st, int mstep) {
  HostPlan h;
  int rc = get_plan(ny, ny, 0, 0, h);
  const int nh = nx / 2 + 1;
  if (rc || (rc = grid_check((long long)B * 2 * nh))) return rc;
  const T s = (T)(1.0 / ((double)nx * (double)ny));
  hipLaunchKernelGGL(k_czt_col_apply<T>, dim3(B * 2 * nh), dim3(block_threads(h.L)), smem_bytes<T>(h.L), st, W, nx,
                     nh, lamx, lamy, lam2d, ip, nrep, s, mstep, dev_plan<T>(h), twiddles<T>());
  return check_launch("czt_col_apply");
}

#define MJX_CZT_DEF(T)                                                                                        \
  int czt_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,     \
                       cpx<T>* W, hipStream_t st, double in_scale) {                                          \
    return row_fwd_pack_t<T>(nx, B, ny, a, b, pitch, istride, W, st, in_scale);                               \
  }                                                                                                           \
  int czt_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,                \
                         long long istride, int ghosts, hipStream_t st, double keep, const InstParams* react) { \
    return row_inv_unpack_t<T>(nx, B, ny, W, a, b, pitch, istride, ghosts, st, keep, react);                  \
  }                                                                                                           \
  int czt_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) { \
    return row_gs_t<T>(nx, B, ny, W, ip, st, rxn, mstep);                                                     \
  }                                                                                                           \
  int czt_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,          \
                    const InstParams* ip, int nrep, hipStream_t st, int mstep) {                              \
    return col_apply_t<T>(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep);                              \
  }

MJX_CZT_DEF(float)
MJX_CZT_DEF(double)
#undef MJX_CZT_DEF

// ------------------------------------------------------------ non-periodic
namespace {

enum { TF_DST1 = 0, TF_DCT2, TF_DCT3, TF_DST2, TF_DST3 };

// Period, doubled offsets and extracted part of each real transform.
void tf_params(int tf, int n, int& P, int& K2, int& N2, int& part) {
  switch (tf) {
    case TF_DST1: P = 2 * (n + 1); K2 = 2; N2 = 2; part = CZT_NEG_IM; break;
    case TF_DCT2: P = 2 * n; K2 = 0; N2 = 1; part = CZT_RE; break;
    case TF_DCT3: P = 2 * n; K2 = 1; N2 = 0; part = CZT_RE; break;
    case TF_DST2: P = 2 * n; K2 = 2; N2 = 1; part = CZT_NEG_IM; break;
    default: P = 2 * n; K2 = 1; N2 = 2; part = CZT_NEG_IM; break;
  }
}

int fwd_tf(int kind) { return kind == R2R_DST ? TF_DST1 : (kind == R2R_DCT ? TF_DCT2 : TF_DST2); }
int inv_tf(int kind) { return kind == R2R_DST ? TF_DST1 : (kind == R2R_DCT ? TF_DCT3 : TF_DST3); }

template <typename T>
int r2r_t(const CztR2RCall& c, hipStream_t st) {
  const int tf1 = c.mode == R2R_INV ? inv_tf(c.kind) : fwd_tf(c.kind);
  int P, K2, N2, part1, part2 = CZT_RE;
  tf_params(tf1, c.n, P, K2, N2, part1);
  HostPlan h1, h2;
  int rc = get_plan(c.n, P, K2, N2, h1);
  if (rc) return rc;
  if (c.mode == R2R_APPLY) {
    tf_params(inv_tf(c.kind), c.n, P, K2, N2, part2);
    if ((rc = get_plan(c.n, P, K2, N2, h2))) return rc;
  } else {
    h2 = h1;
  }
  if ((rc = grid_check((long long)c.B * c.nlines))) return rc;
  CztR2RArgs a{};
  a.in = c.in;
  a.in_es = c.in_es;
  a.in_ls = c.in_ls;
  a.in_is = c.in_is;
  a.out = c.out;
  a.out_es = c.out_es;
  a.out_ls = c.out_ls;
  a.out_is = c.out_is;
  a.nlines = c.nlines;
  a.part1 = part1;
  a.part2 = part2;
  a.sc1 = c.mode == R2R_APPLY ? 1.0 : c.scale;
  a.sc2 = 1.0;
  a.kind = c.kind;
  a.accumulate = c.accumulate;
  a.lam_line = c.lam_line;
  a.lam_cross = c.lam_cross;
  a.n_cross = c.n_cross;
  a.inst = c.inst;
  a.mmode = c.mmode;
  a.beta0 = c.beta0;
  a.lift_on = c.lift ? 1 : 0;
  for (int i = 0; i < 4; ++i) a.lift[i] = c.lift ? c.lift[i] : 0.0;
  const dim3 grid(c.B * c.nlines), block(block_threads(h1.L));
  if (c.mode == R2R_APPLY)
    hipLaunchKernelGGL((k_czt_r2r<T, true>), grid, block, smem_bytes<T>(h1.L), st, a, dev_plan<T>(h1),
                       dev_plan<T>(h2), twiddles<T>());
  else
    hipLaunchKernelGGL((k_czt_r2r<T, false>), grid, block, smem_bytes<T>(h1.L), st, a, dev_plan<T>(h1),
                       dev_plan<T>(h2), twiddles<T>());
  return check_launch("czt_r2r");
}

}  // namespace

int czt_r2r(const CztR2RCall& c, hipStream_t st) {
  if (!czt_len_ok(c.n)) return fail(MJX_E_SIZE, "transform length " + std::to_string(c.n) + " not in [2, 2048]");
  if (c.dtype == MJX_F32) return r2r_t<float>(c, st);
  return r2r_t<double>(c, st);
}

}  // namespace mjx
