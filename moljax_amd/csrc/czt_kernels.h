// Any-length line transforms by the chirp-z (Bluestein) algorithm, for grid
// sizes the power-of-two Stockham kernels do not take directly.
//
// Every transform here is a generalised DFT with half-integer offsets,
//   y_k = sum_{n<N} x_n exp(-2 pi i (k + ko)(n + no) / P),   k < N,
// which with 2 K M = K^2 + M^2 - (K - M)^2 (K = k + ko, M = n + no) becomes
//   y_k = c_k * sum_n (x_n d_n) e_{k-n},
//   c_k = exp(-i pi (k+ko)^2 / P),  d_n = exp(-i pi (n+no)^2 / P),
//   e_m = exp(+i pi (m + ko - no)^2 / P),  |m| < N,
// a linear convolution of length 2N-1, done as a circular one of a
// power-of-two length L >= 2N - 1: the Stockham line FFT of mjx_fft.h
// forward, a pointwise product with FFT_L(e)/L, and the line FFT inverse,
// all on the registers of the threads that hold the line (thread j keeps the
// points j + TN m on input and output, so the chirp factors, the product and
// a multiplier between two transforms need no data movement).  The tables c,
// d and FFT_L(e)/L are evaluated in FP64 on the host (exact integer reduction
// of the squared offsets modulo 8P) and rounded once.
//
// The real transforms of the non-periodic solves, with N points per line:
//   DST-I   X_k = sum x_n sin(pi (k+1)(n+1)/(N+1))  P = 2(N+1), ko = no = 1, -Im
//   DCT-II  X_k = sum x_n cos(pi k (2n+1)/(2N))      P = 2N, ko = 0, no = 1/2, Re
//   DCT-III x_n = sum X_k cos(pi k (2n+1)/(2N))      P = 2N, ko = 1/2, no = 0, Re
//   DST-II  X_k = sum x_n sin(pi (k+1)(2n+1)/(2N))   P = 2N, ko = 1, no = 1/2, -Im
//   DST-III x_n = sum X_k sin(pi (k+1)(2n+1)/(2N))   P = 2N, ko = 1/2, no = 1, -Im
// (output index k carries ko, input index n carries no).  The periodic
// complex DFT is P = N, ko = no = 0; its inverse is conj(DFT(conj(x))).
//
// Tiles: a workgroup holds NL lines of TN = L/P threads each.  Row tiles
// (lines contiguous in memory) keep one padded LDS line per line; column
// tiles (NL adjacent lines, elements strided) interleave the NL lines in LDS
// and give consecutive threads consecutive lines, so loads and stores of a
// tile row are contiguous.  Workgroups are remapped per XCD (xcd_remap).
#pragma once
#include <hip/hip_runtime.h>

#include "etd_kernels.h"
#include "mjx_common.h"
#include "mjx_fft.h"
#include "r2r_kernels.h"
#include "spectral_api.h"
#include "spectral_kernels.h"

namespace mjx {

// Device tables of one chirp-z plan.
template <typename T>
struct CztPlan {
  const cpx<T>* c;  // output chirp, n entries
  const cpx<T>* d;  // input chirp, n entries
  const cpx<T>* e;  // FFT_L of the convolution kernel, times 1/L
  int n;
};

enum { CZT_ROW = 0, CZT_COL = 1 };

template <typename T, int L>
struct CztCfg {
  static constexpr int P = plan_p(L);
  static constexpr int TN = L / P;
  static constexpr int BT0 = sizeof(T) == 8 ? 128 : 256;
  static constexpr int NL = TN >= BT0 ? 1 : BT0 / TN;  // lines per workgroup
  static constexpr int BT = NL * TN;
  static constexpr int LP = lds_line_len(L, 1);        // padded row line
  static constexpr int LDS = NL * LP;                  // >= NL * L (column tiles)
};

// Thread placement of a tile: line within the tile and lane j in the line.
template <int ORIENT, int TN, int NL>
__device__ __forceinline__ void czt_place(int& line, int& j) {
  if constexpr (ORIENT == CZT_ROW) {
    line = threadIdx.x / TN;
    j = threadIdx.x % TN;
  } else {
    line = threadIdx.x % NL;
    j = threadIdx.x / NL;
  }
}

// This line's LDS base in a tile.
template <typename T, int L, int ORIENT>
__device__ __forceinline__ cpx<T>* czt_lline(cpx<T>* lds, int line) {
  return ORIENT == CZT_ROW ? lds + line * lds_line_len(L, 1) : lds + line;
}

// The circular convolution of the line in v (x_n d_n at natural index
// j + TN m, zero past n) with e; v then holds the convolution at the same
// indices.  Every thread of the workgroup calls it (barriers).
template <typename T, int L, int ORIENT, int NL>
__device__ __forceinline__ void czt_conv(cpx<T> (&v)[plan_p(L)], int j, cpx<T>* lline, const TwCache<T, L>& twc,
                                         const cpx<T>* __restrict__ e) {
  constexpr int TN = L / plan_p(L);
  constexpr int ES = ORIENT == CZT_ROW ? 1 : NL;
  constexpr int PAD = ORIENT == CZT_ROW ? 1 : 0;
  line_fft<T, L, -1, ES, PAD>(v, j, lline, twc);
#pragma unroll
  for (int m = 0; m < plan_p(L); ++m) v[m] = cmul(v[m], e[j + TN * m]);
  line_fft<T, L, +1, ES, PAD>(v, j, lline, twc);
}

// ------------------------------------------------------ real line transforms
enum { CZT_RE = 0, CZT_NEG_IM = 1 };

struct CztR2RArgs {
  const void* in;
  long long in_es, in_ls, in_is;
  void* out;
  long long out_es, out_ls, out_is;
  int nlines;
  int tiles;            // line tiles per instance
  int part1, part2;     // CZT_RE or CZT_NEG_IM of the first / second transform
  double sc1, sc2;      // output scales
  int kind;             // 0 DST-I, 1 DCT-II, 2 DST-II (normalisation of the apply)
  int accumulate;
  const void* lam_line;
  const void* lam_cross;
  int n_cross;          // cross-dimension length (its normalisation), 0 = none
  const double* inst;
  int mmode;
  double beta0;
  int lift_on;
  double lift[4];
};

// Inverse-pair normalisation of coefficient k of an n-point transform:
// DST-I 2/(n+1); DCT-II/III 1/n at k = 0, else 2/n; DST-II/III 1/n at
// k = n-1, else 2/n.
__device__ __forceinline__ double czt_norm(int kind, int k, int n) {
  if (kind == 0) return 2.0 / (n + 1);
  if (kind == 1) return (k == 0 ? 1.0 : 2.0) / n;
  return (k == n - 1 ? 1.0 : 2.0) / n;
}

// APPLY: forward transform, spectral multiplier, inverse transform (second
// plan); otherwise one transform times sc1.
// NLW lines per workgroup (0: CztCfg's default).
template <typename T, int L, int NLW>
struct CztR2RCfg {
  static constexpr int NL = NLW > 0 ? NLW : CztCfg<T, L>::NL;
  static constexpr int BT = NL * CztCfg<T, L>::TN;
  static constexpr int LDS = NL * CztCfg<T, L>::LP;
};

template <typename T, int L, int ORIENT, bool APPLY, int NLW = 0>
__global__ __launch_bounds__((CztR2RCfg<T, L, NLW>::BT)) void k_czt_r2r(CztR2RArgs a, CztPlan<T> p1, CztPlan<T> p2,
                                                                      const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  constexpr int P = C::P, TN = C::TN, NL = CztR2RCfg<T, L, NLW>::NL;
  __shared__ cpx<T> lds[CztR2RCfg<T, L, NLW>::LDS];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / a.tiles;
  int tl, j;
  czt_place<ORIENT, TN, NL>(tl, j);
  const int line = (w - inst * a.tiles) * NL + tl;
  const bool valid = line < a.nlines;
  const int N = p1.n;
  const T* in = (const T*)a.in + inst * a.in_is + (long long)line * a.in_ls;
  T* out = (T*)a.out + inst * a.out_is + (long long)line * a.out_ls;
  double lw = 0.0;
  if (a.lift_on) lw = a.mmode == R2R_POISSON ? -1.0 : a.inst[8 * inst];
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int n = j + TN * m;
    v[m] = mk((T)0, (T)0);
    if (valid && n < N) {
      T x = in[(long long)n * a.in_es];
      if (a.lift_on) {
        const double add = (n == 0 ? a.lift[0] : 0.0) + (n == N - 1 ? a.lift[1] : 0.0) +
                           (line == 0 ? a.lift[2] : 0.0) + (line == a.nlines - 1 ? a.lift[3] : 0.0);
        x += (T)(lw * add);
      }
      v[m] = scale(p1.d[n], x);
    }
  }
  cpx<T>* ll = czt_lline<T, L, ORIENT>(lds, tl);
  czt_conv<T, L, ORIENT, NL>(v, j, ll, twc, p1.e);
  if constexpr (APPLY) {
    const double* ip = a.inst ? a.inst + 8 * inst : nullptr;
    const T beta = ip ? (T)ip[0] : (T)a.beta0, alpha = ip ? (T)ip[2] : (T)0;
    const int power = (ip && a.mmode == R2R_HELM_POW) ? (int)ip[7] : 1;
    const T lc = (a.lam_cross && valid) ? ((const T*)a.lam_cross)[line] : (T)0;
    const double fc = a.n_cross > 0 ? czt_norm(a.kind, line, a.n_cross) : 1.0;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int k = j + TN * m;
      cpx<T> z = mk((T)0, (T)0);
      if (k < N) {
        const cpx<T> y = cmul(p1.c[k], v[m]);
        const T x = a.part1 == CZT_RE ? y.re : -y.im;
        const T mu = r2r_mult<T>(((const T*)a.lam_line)[k] + lc, alpha, beta, a.mmode, power) *
                     (T)(czt_norm(a.kind, k, N) * fc);
        z = scale(p2.d[k], x * mu);
      }
      v[m] = z;
    }
    czt_conv<T, L, ORIENT, NL>(v, j, ll, twc, p2.e);
  }
  if (!valid) return;
  const cpx<T>* co = APPLY ? p2.c : p1.c;
  const int part = APPLY ? a.part2 : a.part1;
  const T sc = (T)(APPLY ? a.sc2 : a.sc1);
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int k = j + TN * m;
    if (k < N) {
      const cpx<T> y = cmul(co[k], v[m]);
      const T r = (part == CZT_RE ? y.re : -y.im) * sc;
      T* o = out + (long long)k * a.out_es;
      *o = a.accumulate ? *o + r : r;
    }
  }
}

// ------------------------------------------------------ periodic spectral path
// Layout of W for an ny x nx instance: rows of 2*nh complex values
// (nh = nx/2 + 1), the x-half-spectrum A_kx of field a at [kx] and B_kx of
// field b at [nh + kx], kx = 0..nh-1; A_{-kx} = conj(A_kx).

// Z_k = A_k + i B_k of the packed line a + i b from the half-spectra.
template <typename T>
__device__ __forceinline__ cpx<T> czt_merge(const cpx<T>* __restrict__ wrow, int k, int nx, int nh) {
  if (k < nh) {
    const cpx<T> A = wrow[k], Bv = wrow[nh + k];
    return mk(A.re - Bv.im, A.im + Bv.re);
  }
  const cpx<T> A = wrow[nx - k], Bv = wrow[nh + nx - k];
  return mk(A.re + Bv.im, Bv.re - A.im);
}

// Half-spectra of the two real lines packed in the spectrum Z (the row
// transform's output, staged in the line's LDS in natural order).
template <typename T, int L>
__device__ __forceinline__ void czt_split_store(cpx<T> (&v)[plan_p(L)], const CztPlan<T>& p, int j, cpx<T>* lrow,
                                                cpx<T>* __restrict__ wrow, int nh, bool valid) {
  constexpr int TN = L / plan_p(L);
  const int nx = p.n;
#pragma unroll
  for (int m = 0; m < plan_p(L); ++m) {
    const int k = j + TN * m;
    if (k < nx) lrow[lds_addr<1, 1>(k)] = cmul(p.c[k], v[m]);
  }
  __syncthreads();
  if (!valid) return;
  for (int k = j; k < nh; k += TN) {
    const int km = k == 0 ? 0 : nx - k;
    const cpx<T> z = lrow[lds_addr<1, 1>(k)], zm = conjg(lrow[lds_addr<1, 1>(km)]);
    wrow[k] = mk((T)0.5 * (z.re + zm.re), (T)0.5 * (z.im + zm.im));
    wrow[nh + k] = mk((T)0.5 * (z.im - zm.im), (T)0.5 * (zm.re - z.re));
  }
}

template <typename T, int L>
__global__ __launch_bounds__((CztCfg<T, L>::BT)) void k_czt_row_fwd_pack(
    const T* __restrict__ a, const T* __restrict__ b, long long pitch, long long istride, cpx<T>* __restrict__ W,
    int ny, int nh, int tiles, T in_scale, CztPlan<T> px, const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  __shared__ cpx<T> lds[C::LDS];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / tiles;
  int rl, j;
  czt_place<CZT_ROW, C::TN, C::NL>(rl, j);
  const int y = (w - inst * tiles) * C::NL + rl;
  const bool valid = y < ny;
  const T* ra = a + inst * istride + (long long)y * pitch;
  const T* rb = b ? b + inst * istride + (long long)y * pitch : nullptr;
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int n = j + C::TN * m;
    v[m] = (valid && n < px.n) ? cmul(px.d[n], mk(ra[n] * in_scale, rb ? rb[n] * in_scale : (T)0))
                               : mk((T)0, (T)0);
  }
  cpx<T>* lrow = czt_lline<T, L, CZT_ROW>(lds, rl);
  czt_conv<T, L, CZT_ROW, C::NL>(v, j, lrow, twc, px.e);
  czt_split_store<T, L>(v, px, j, lrow, W + ((long long)inst * ny + y) * 2 * nh, nh, valid);
}

template <typename T, int L>
__global__ __launch_bounds__((CztCfg<T, L>::BT)) void k_czt_row_inv_unpack(
    const cpx<T>* __restrict__ W, T* __restrict__ a, T* __restrict__ b, long long pitch, long long istride, int ny,
    int nh, int tiles, int ghosts, T keep_w, const InstParams* __restrict__ react, CztPlan<T> px,
    const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  __shared__ cpx<T> lds[C::LDS];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / tiles;
  int rl, j;
  czt_place<CZT_ROW, C::TN, C::NL>(rl, j);
  const int y = (w - inst * tiles) * C::NL + rl;
  const bool valid = y < ny;
  const int nx = px.n;
  const cpx<T>* wrow = W + ((long long)inst * ny + y) * 2 * nh;
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int k = j + C::TN * m;
    v[m] = (valid && k < nx) ? cmul(px.d[k], conjg(czt_merge(wrow, k, nx, nh))) : mk((T)0, (T)0);
  }
  czt_conv<T, L, CZT_ROW, C::NL>(v, j, czt_lline<T, L, CZT_ROW>(lds, rl), twc, px.e);
  if (!valid) return;
  T* ra = a + inst * istride + (long long)y * pitch;
  T* rb = b ? b + inst * istride + (long long)y * pitch : nullptr;
  const long long offs[3] = {0, (y == 0) ? (long long)ny * pitch : 0, (y == ny - 1) ? -(long long)ny * pitch : 0};
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int x = j + C::TN * m;
    if (x >= nx) continue;
    cpx<T> z = conjg(cmul(px.c[x], v[m]));
    if (react) {
      const T F = (T)react[inst].F, kk = (T)react[inst].k, hdt = (T)(0.5 * react[inst].dt);
      T ru, rv;
      gs_react(z.re, z.im, F, kk, ru, rv);
      z = mk(z.re + hdt * ru, z.im + hdt * rv);
    }
    if (keep_w != (T)0) {
      const T fresh = (T)1 - keep_w;
      z.re = keep_w * ra[x] + fresh * z.re;
      if (rb) z.im = keep_w * rb[x] + fresh * z.im;
    }
    ra[x] = z.re;
    if (rb) rb[x] = z.im;
    if (ghosts) {
      // periodic ghost ring, x first then y (moljax/core/bc.py:202-215)
      for (int o = 0; o < 3; ++o) {
        if (o > 0 && offs[o] == 0) continue;
        T* qa = ra + offs[o];
        T* qb = rb ? rb + offs[o] : nullptr;
        if (o > 0) {
          qa[x] = z.re;
          if (qb) qb[x] = z.im;
        }
        if (x == 0) {
          qa[nx] = z.re;
          if (qb) qb[nx] = z.im;
        }
        if (x == nx - 1) {
          qa[-1] = z.re;
          if (qb) qb[-1] = z.im;
        }
      }
    }
  }
}

// Fused Gray-Scott row pass on the half-spectra: inverse x-DFT, a
// pointwise step, forward x-DFT, Win -> Wout (may alias).  NLOP false: the
// explicit reaction in place (RXN_HEUN, or forward Euler over dt, dt/2 or
// two dt/2 steps); NLOP true: the nonlinear term N(u, v) / nx of the ETD
// chains (their x-spectra carry the x-DFT divided by nx).
template <typename T, int L, bool NLOP>
__global__ __launch_bounds__((CztCfg<T, L>::BT)) void k_czt_row_gs(const cpx<T>* Win, cpx<T>* Wout,
                                                                 const InstParams* __restrict__ ip, int ny, int nh,
                                                                 int tiles, int rxn, int mstep, CztPlan<T> px,
                                                                 const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  __shared__ cpx<T> lds[C::LDS];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / tiles;
  if (mstep >= 0 && mstep >= (int)ip[inst].aux) return;  // the whole workgroup is one instance
  int rl, j;
  czt_place<CZT_ROW, C::TN, C::NL>(rl, j);
  const int y = (w - inst * tiles) * C::NL + rl;
  const bool valid = y < ny;
  const int nx = px.n;
  const cpx<T>* wrow = Win + ((long long)inst * ny + y) * 2 * nh;
  const T F = (T)ip[inst].F, kk = (T)ip[inst].k, dt = (T)ip[inst].dt, hdt = (T)(0.5 * ip[inst].dt);
  const T inv_nx = (T)(1.0 / nx);
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int k = j + C::TN * m;
    v[m] = (valid && k < nx) ? cmul(px.d[k], conjg(czt_merge(wrow, k, nx, nh))) : mk((T)0, (T)0);
  }
  cpx<T>* lrow = czt_lline<T, L, CZT_ROW>(lds, rl);
  czt_conv<T, L, CZT_ROW, C::NL>(v, j, lrow, twc, px.e);
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int x = j + C::TN * m;
    cpx<T> o = mk((T)0, (T)0);
    if (valid && x < nx) {
      const cpx<T> z = conjg(cmul(px.c[x], v[m]));
      const T u0 = z.re, v0 = z.im;
      T r1u, r1v, r2u, r2v;
      gs_react(u0, v0, F, kk, r1u, r1v);
      if constexpr (NLOP) {
        o = mk(r1u * inv_nx, r1v * inv_nx);
      } else if (rxn == RXN_HEUN) {
        const T ut = u0 + dt * r1u, vt = v0 + dt * r1v;
        gs_react(ut, vt, F, kk, r2u, r2v);
        o = mk(u0 + hdt * (r1u + r2u), v0 + hdt * (r1v + r2v));
      } else if (rxn == RXN_EULER) {
        o = mk(u0 + dt * r1u, v0 + dt * r1v);
      } else {
        T ut = u0 + hdt * r1u, vt = v0 + hdt * r1v;
        if (rxn == RXN_HALF2) {
          gs_react(ut, vt, F, kk, r2u, r2v);
          ut = ut + hdt * r2u;
          vt = vt + hdt * r2v;
        }
        o = mk(ut, vt);
      }
      o = cmul(px.d[x], o);
    }
    v[m] = o;
  }
  czt_conv<T, L, CZT_ROW, C::NL>(v, j, lrow, twc, px.e);
  czt_split_store<T, L>(v, px, j, lrow, Wout + ((long long)inst * ny + y) * 2 * nh, nh, valid);
}

// Column pass over the 2*nh half-spectrum columns: y-DFT, multiplier
// spec_mult(lam, alpha_f, beta_f, nrep) * s, inverse y-DFT, in place.  The
// per-instance step rules are those of the power-of-two column pass.
template <typename T, int L>
__global__ __launch_bounds__((CztCfg<T, L>::BT)) void k_czt_col_apply(
    cpx<T>* __restrict__ W, int nx, int nh, int tiles, const T* __restrict__ lamx, const T* __restrict__ lamy,
    const T* __restrict__ lam2d, const InstParams* __restrict__ ip, int nrep_arg, T s, int mstep, CztPlan<T> py,
    const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  __shared__ cpx<T> lds[C::LDS];
  const int w = xcd_remap_rev(blockIdx.x, gridDim.x);
  const int inst = w / tiles;
  int nrep = nrep_arg;
  if (mstep >= 0 || nrep_arg == NREP_PER_INSTANCE || mstep == MSTEP_ACTIVE) {
    const int ni = (int)ip[inst].aux;
    if (mstep == MSTEP_ACTIVE) {
      if (ni <= 0) return;
    } else if (mstep >= 0) {
      if (mstep >= ni) return;
      nrep = mstep == ni - 1 ? 1 : 2;
    } else {
      nrep = ni;
    }
  }
  int cl, j;
  czt_place<CZT_COL, C::TN, C::NL>(cl, j);
  const int cols = 2 * nh;
  const int c = (w - inst * tiles) * C::NL + cl;
  const bool valid = c < cols;
  const int ny = py.n;
  const bool fb = c >= nh;
  const int kx = valid ? (fb ? c - nh : c) : 0;
  const T alpha = (T)(fb ? ip[inst].alpha_b : ip[inst].alpha_a);
  const T beta = (T)(fb ? ip[inst].beta_b : ip[inst].beta_a);
  cpx<T>* col = W + (long long)inst * ny * cols + c;
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    v[m] = (valid && ky < ny) ? cmul(py.d[ky], col[(long long)ky * cols]) : mk((T)0, (T)0);
  }
  cpx<T>* lc = czt_lline<T, L, CZT_COL>(lds, cl);
  czt_conv<T, L, CZT_COL, C::NL>(v, j, lc, twc, py.e);
  const T lx = (lam2d || !valid) ? (T)0 : lamx[kx];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    cpx<T> z = mk((T)0, (T)0);
    if (valid && ky < ny) {
      const T lam = lam2d ? lam2d[(long long)ky * nx + kx] : lx + lamy[ky];
      z = cmul(py.d[ky], conjg(scale(cmul(py.c[ky], v[m]), spec_mult<T, 0>(lam, alpha, beta, nrep, s))));
    }
    v[m] = z;
  }
  czt_conv<T, L, CZT_COL, C::NL>(v, j, lc, twc, py.e);
  if (!valid) return;
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    if (ky < ny) col[(long long)ky * cols] = conjg(cmul(py.c[ky], v[m]));
  }
}

// ETD column stage over the half-spectrum columns: y-DFT of the stage
// input(s), the stage arithmetic of etd_stage on the separated planes (row
// stride sw; column c of the half-spectra is separated column c), inverse
// y-DFT into Wout.  Every column is one field at one kx.
template <typename T, int L, int STAGE>
__global__ __launch_bounds__((CztCfg<T, L>::BT)) void k_czt_col_etd(
    const cpx<T>* __restrict__ Win, const cpx<T>* __restrict__ Win2, cpx<T>* __restrict__ Uh, cpx<T>* __restrict__ Nn,
    cpx<T>* __restrict__ Fh, const T* __restrict__ coef, cpx<T>* __restrict__ Wout, int nh, int sw, int tiles,
    CztPlan<T> py, const cpx<T>* __restrict__ tw) {
  using C = CztCfg<T, L>;
  __shared__ cpx<T> lds[C::LDS];
  const int w = xcd_remap_rev(blockIdx.x, gridDim.x);
  const int inst = w / tiles;
  int cl, j;
  czt_place<CZT_COL, C::TN, C::NL>(cl, j);
  const int cols = 2 * nh;
  const int c = (w - inst * tiles) * C::NL + cl;
  const bool valid = c < cols;
  const int ny = py.n;
  const long long tstride = (long long)ny * sw;
  const T s = (T)(1.0 / (double)ny);
  const long long woff = (long long)inst * ny * cols + c;
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P], v2[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    const bool in = valid && ky < ny;
    v[m] = in ? cmul(py.d[ky], Win[woff + (long long)ky * cols]) : mk((T)0, (T)0);
    v2[m] = (two_inputs<STAGE>() && in) ? cmul(py.d[ky], Win2[woff + (long long)ky * cols]) : mk((T)0, (T)0);
  }
  cpx<T>* lc = czt_lline<T, L, CZT_COL>(lds, cl);
  czt_conv<T, L, CZT_COL, C::NL>(v, j, lc, twc, py.e);
  if constexpr (two_inputs<STAGE>()) czt_conv<T, L, CZT_COL, C::NL>(v2, j, lc, twc, py.e);
  const long long sbase = (long long)inst * tstride;
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    cpx<T> z = mk((T)0, (T)0);
    if (valid && ky < ny) {
      const cpx<T> x = cmul(py.c[ky], v[m]);
      const cpx<T> x2 = two_inputs<STAGE>() ? cmul(py.c[ky], v2[m]) : mk((T)0, (T)0);
      const long long crow = (long long)ky * sw + c;
      z = cmul(py.d[ky], conjg(etd_stage<T, STAGE>(sbase + crow, crow, tstride, coef, x, x2, Uh, Nn, Fh, s)));
    }
    v[m] = z;
  }
  if constexpr (STAGE == STAGE_STORE_N) return;
  czt_conv<T, L, CZT_COL, C::NL>(v, j, lc, twc, py.e);
  if (!valid) return;
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int ky = j + C::TN * m;
    if (ky < ny) Wout[woff + (long long)ky * cols] = conjg(cmul(py.c[ky], v[m]));
  }
}

#define MJX_FOR_CZT_LENGTHS(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

}  // namespace mjx
