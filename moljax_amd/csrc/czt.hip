// Host side of the any-length (chirp-z) path: the plan tables (evaluated in
// FP64 on the host, one set per device and plan) and the argument set-up of
// the non-periodic line transforms.  Kernel launchers: czt_launch.h.
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

std::mutex g_mu;
std::map<std::array<int, 5>, CztTables> g_plans;

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

}  // namespace

int czt_tables(int n, int P, int K2, int N2, CztTables& out) {
  if (!czt_len_ok(n)) return fail(MJX_E_SIZE, "chirp-z length " + std::to_string(n) + " not in [2, 2048]");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(MJX_E_HIP, "hipGetDevice failed");
  const std::array<int, 5> key{dev, n, P, K2, N2};
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    out = it->second;
    return MJX_OK;
  }
  CztTables h;
  h.n = n;
  h.L = 16;
  while (h.L < 2 * n - 1) h.L <<= 1;
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
  cpx<float>* d32 = nullptr;
  cpx<double>* d64 = nullptr;
  if (hipMalloc(&d32, cnt * sizeof(cpx<float>)) != hipSuccess ||
      hipMalloc(&d64, cnt * sizeof(cpx<double>)) != hipSuccess)
    return fail(MJX_E_HIP, "chirp-z table allocation failed");
  if (hipMemcpy(d32, t32.data(), cnt * sizeof(cpx<float>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d64, t64.data(), cnt * sizeof(cpx<double>), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MJX_E_HIP, "chirp-z table upload failed");
  h.t32 = d32;
  h.t64 = d64;
  g_plans[key] = h;
  out = h;
  return MJX_OK;
}

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

}  // namespace

int czt_prepare_r2r(int kind, int n) {
  for (int tf : {fwd_tf(kind), inv_tf(kind)}) {
    int P, K2, N2, part;
    tf_params(tf, n, P, K2, N2, part);
    CztTables h;
    const int rc = czt_tables(n, P, K2, N2, h);
    if (rc) return rc;
  }
  return MJX_OK;
}

int czt_prepare_periodic(int n) {
  CztTables h;
  return czt_tables(n, n, 0, 0, h);
}

int czt_r2r(const CztR2RCall& c, hipStream_t st) {
  const int tf1 = c.mode == R2R_INV ? inv_tf(c.kind) : fwd_tf(c.kind);
  int P, K2, N2, part1, part2 = CZT_RE;
  tf_params(tf1, c.n, P, K2, N2, part1);
  CztTables h1, h2;
  int rc = czt_tables(c.n, P, K2, N2, h1);
  if (rc) return rc;
  const bool apply = c.mode == R2R_APPLY;
  if (apply) {
    tf_params(inv_tf(c.kind), c.n, P, K2, N2, part2);
    if ((rc = czt_tables(c.n, P, K2, N2, h2))) return rc;
  } else {
    h2 = h1;
  }
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
  a.sc1 = apply ? 1.0 : c.scale;
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
  // lines with contiguous elements take row tiles, strided lines column tiles
  const int orient = c.in_es == 1 ? CZT_ROW : CZT_COL;
  if (c.dtype == MJX_F32) return czt_r2r_launch(a, orient, apply, h1, h2, c.B, st, 0.0f);
  return czt_r2r_launch(a, orient, apply, h1, h2, c.B, st, 0.0);
}

}  // namespace mjx
