// Periodic 2-D spectral kernels and their launchers, included once per
// precision by spectral_f32.hip / spectral_f64.hip (explicit instantiation of
// Launch<T>) so the two precisions compile in parallel.
// Kernels: row passes (x-FFT with two-real-field
// packing), column pass (y-FFT, spectral multiply, inverse y-FFT) and the
// fused Gray-Scott row pass (inverse x-FFT, Heun reaction, forward x-FFT).
//
// Packed x-spectral layout (one complex row of nx values per grid row), for
// the real fields a and b with row spectra A(kx), B(kx) (Hermitian in kx):
//   col 0        A(0) + i*A(nx/2)        (both real)
//   col nx/2     B(0) + i*B(nx/2)        (both real)
//   col p, 0<p<nx/2          A(p)
//   col nx/2+p, 0<p<nx/2     B(p)
// so that every column except 0 and nx/2 carries one field at one kx and the
// column pass is a plain complex y-FFT with a real multiplier.  Columns 0 and
// nx/2 carry two real sequences with different multipliers ma, mb; after the
// y-FFT, out(ky) = (ma+mb)/2 Z(ky) + (ma-mb)/2 conj(Z(-ky)).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "mjx_common.h"
#include "mjx_fft.h"
#include "spectral_api.h"

namespace mjx {

// Current device's twiddle tables (runtime.hip).
const cpx<float>* twiddles_f32();
const cpx<double>* twiddles_f64();
template <typename T>
inline const cpx<T>* twiddles() {
  if constexpr (std::is_same<T, float>::value) return twiddles_f32();
  else return twiddles_f64();
}

// ------------------------------------------------------------ pack/unpack
// Row LDS layout: one padded line per row (lds_addr<1, 1>).
template <typename T>
__device__ __forceinline__ cpx<T> lrd(const cpx<T>* l, int idx) { return l[lds_addr<1, 1>(idx)]; }
template <typename T>
__device__ __forceinline__ void lwr(cpx<T>* l, int idx, cpx<T> v) { l[lds_addr<1, 1>(idx)] = v; }

// Natural-order complex row spectrum Z of a + i*b in LDS; returns the packed
// value for column p.
template <typename T, int NX>
__device__ __forceinline__ cpx<T> pack_at(const cpx<T>* l, int p) {
  constexpr int H = NX / 2;
  if (p == 0) return mk(lrd(l, 0).re, lrd(l, H).re);
  if (p == H) return mk(lrd(l, 0).im, lrd(l, H).im);
  if (p < H) {
    const cpx<T> z = lrd(l, p), zm = lrd(l, NX - p);
    return mk((z.re + zm.re) * (T)0.5, (z.im - zm.im) * (T)0.5);
  }
  const int e = p - H;
  const cpx<T> z = lrd(l, e), zm = lrd(l, NX - e);
  // (z - conj(zm)) / (2i)
  const T dre = z.re - zm.re, dim = z.im + zm.im;
  return mk(dim * (T)0.5, -dre * (T)0.5);
}

// Packed row in LDS; returns Z(e) = A(e) + i*B(e).
template <typename T, int NX>
__device__ __forceinline__ cpx<T> unpack_at(const cpx<T>* l, int e) {
  constexpr int H = NX / 2;
  if (e == 0) return mk(lrd(l, 0).re, lrd(l, H).re);
  if (e == H) return mk(lrd(l, 0).im, lrd(l, H).im);
  cpx<T> A, B;
  if (e < H) {
    A = lrd(l, e);
    B = lrd(l, H + e);
  } else {
    A = conjg(lrd(l, NX - e));
    B = conjg(lrd(l, H + NX - e));
  }
  return mk(A.re - B.im, A.im + B.re);
}

// 1/d: FP64 exact division; FP32 hardware reciprocal refined by one Newton
// step (error below 1 ulp).
template <typename T>
__device__ __forceinline__ T recip(T d) {
  if constexpr (std::is_same<T, double>::value) {
    return 1.0 / d;
  } else {
    float r = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r, 1.0f), r, r);
  }
}

// s * ((1 + alpha*lam) / (1 - beta*lam))^nrep with |1 - beta*lam| clamped
// below at 1e-14 (moljax/core/fft_solvers.py:280-283).  NREP > 0 fixes the
// power at compile time; NREP == 0 takes the runtime nrep, where nrep = -1
// selects s * alpha * lam, nrep = -2 selects s * exp(beta * lam) and
// NREP_INC1 / NREP_INC2 select the increments s * (g - 1) / s * (g^2 - 1) of
// g = (1 + alpha*lam) / (1 - beta*lam);
// NREP == -1 is the s * alpha * lam form fixed at compile time.
template <typename T, int NREP>
__device__ __forceinline__ T spec_mult(T lam, T alpha, T beta, int nrep, T s) {
  if constexpr (NREP == -1) return alpha * lam * s;  // the matvec form, without the runtime dispatch
  T den = (T)1 - beta * lam;
  if (fabs(den) < (T)1e-14) den = (T)1e-14;
  if constexpr (NREP == NREP_INC1 || NREP == NREP_INC2) {  // the increment forms at compile time
    T g1;
    if constexpr (std::is_same<T, double>::value) g1 = ((alpha + beta) * lam) / den;
    else g1 = ((alpha + beta) * lam) * recip(den);
    return (NREP == NREP_INC1 ? g1 : g1 * (g1 + (T)2)) * s;
  }
  const T num = (T)1 + alpha * lam;
  T g;
  if constexpr (std::is_same<T, double>::value) g = num / den;
  else g = num * recip(den);
  if constexpr (NREP == 1) {
    return g * s;
  } else if constexpr (NREP == 2) {
    return (g * g) * s;
  } else {
    if (nrep == -1) return alpha * lam * s;
    if (nrep == -2) return exp(beta * lam) * s;
    if (nrep == NREP_INC1 || nrep == NREP_INC2) {
      // increments g - 1 = (alpha + beta) lam / den and g^2 - 1 = (g - 1)(g - 1 + 2),
      // formed without the cancellation of g - 1 near lam = 0
      T g1;
      if constexpr (std::is_same<T, double>::value) g1 = ((alpha + beta) * lam) / den;
      else g1 = ((alpha + beta) * lam) * recip(den);
      return (nrep == NREP_INC1 ? g1 : g1 * (g1 + (T)2)) * s;
    }
    // the fixed-power forms, so a runtime power of 1 or 2 rounds exactly as
    // the NREP = 1 / 2 instantiations do
    if (nrep == 1) return g * s;
    if (nrep == 2) return (g * g) * s;
    T m = (T)1, b = g;
    int n = nrep;
    while (n > 0) {
      if (n & 1) m *= b;
      b *= b;
      n >>= 1;
    }
    return m * s;
  }
}

// Gray-Scott reaction terms, moljax/core/operators.py:655-667.
template <typename T>
__device__ __forceinline__ void gs_react(T u, T v, T F, T k, T& ru, T& rv) {
  const T uvv = u * v * v;
  ru = -uvv + F * ((T)1 - u);
  rv = uvv - (F + k) * v;
}

// ------------------------------------------------------------- row passes
template <int NX, int TR_>
struct RowCfg {
  static constexpr int P = plan_p(NX);
  static constexpr int TN = NX / P;
  static constexpr int TR = TR_;                       // rows per workgroup
  static constexpr int BT = TR * TN;                   // threads per workgroup
  static constexpr int LP = lds_line_len(NX, 1);       // padded LDS row length
};

// Real fields (a, b) -> packed x-spectra in W.
template <typename T, int NX, int TR>
__global__ __launch_bounds__((RowCfg<NX, TR>::BT)) void k_row_fwd_pack(
    const T* __restrict__ a, const T* __restrict__ b, long long pitch, long long istride, cpx<T>* __restrict__ W,
    int ny, int rows_per_inst_tiles, T in_scale, const cpx<T>* __restrict__ tw) {
  using C = RowCfg<NX, TR>;
  __shared__ cpx<T> lds[C::TR * C::LP];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / rows_per_inst_tiles;
  const int tile = w - inst * rows_per_inst_tiles;
  const int j = threadIdx.x % C::TN, rl = threadIdx.x / C::TN;
  const int y = tile * C::TR + rl;
  // the last tile of an instance may hold rows past ny (ny not a multiple of
  // TR): they run the transforms on zeros, for the barriers, and store nothing
  const bool live = y < ny;
  cpx<T>* lrow = lds + rl * C::LP;
  const T* ra = a + inst * istride + (long long)(live ? y : 0) * pitch;
  const T* rb = b ? b + inst * istride + (long long)(live ? y : 0) * pitch : nullptr;
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int x = j + C::TN * m;
    v[m] = live ? mk(ra[x] * in_scale, rb ? rb[x] * in_scale : (T)0) : mk((T)0, (T)0);
  }
  TwCache<T, NX> twc;
  load_twiddles(twc, j, tw);
  line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
#pragma unroll
  for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, v[m]);
  __syncthreads();
  if (!live) return;
  cpx<T>* wrow = W + ((long long)inst * ny + y) * NX;
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int p = j + C::TN * m;
    wrow[p] = pack_at<T, NX>(lrow, p);
  }
}

// Load one packed row into LDS and return the unpacked Z(e) at e = j + TN*m.
template <typename T, int NX, int TR>
__device__ __forceinline__ void load_unpack(const cpx<T>* __restrict__ wrow, bool live, cpx<T>* lrow, int j,
                                            cpx<T> (&v)[RowCfg<NX, TR>::P]) {
  using C = RowCfg<NX, TR>;
  cpx<T> g[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) g[m] = live ? wrow[j + C::TN * m] : mk((T)0, (T)0);
#pragma unroll
  for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, g[m]);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < C::P; ++m) v[m] = unpack_at<T, NX>(lrow, j + C::TN * m);
  __syncthreads();
}

// Packed x-spectra in W -> real fields (a, b), optional periodic ghost ring.
// keep != 0 blends with the destination's previous content:
// out = keep * old + (1 - keep) * new (the averaging stage of IMEX-SSPRK2).
// react != null applies one forward-Euler Gray-Scott step of dt/2 (F, k, dt
// from the instance row) to the physical values before they are written
// (the closing reaction half-step of benchmark_gray_scott.py:268-289).
template <typename T, int NX, int TR>
__global__ __launch_bounds__((RowCfg<NX, TR>::BT)) void k_row_inv_unpack(
    const cpx<T>* __restrict__ W, T* __restrict__ a, T* __restrict__ b, long long pitch, long long istride, int ny,
    int rows_per_inst_tiles, int ghosts, T keep, const cpx<T>* __restrict__ tw,
    const InstParams* __restrict__ react) {
  using C = RowCfg<NX, TR>;
  __shared__ cpx<T> lds[C::TR * C::LP];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / rows_per_inst_tiles;
  const int tile = w - inst * rows_per_inst_tiles;
  const int j = threadIdx.x % C::TN, rl = threadIdx.x / C::TN;
  const int y = tile * C::TR + rl;
  const bool live = y < ny;  // rows past ny: transforms only
  cpx<T>* lrow = lds + rl * C::LP;
  TwCache<T, NX> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
  load_unpack<T, NX, TR>(W + ((long long)inst * ny + (live ? y : 0)) * NX, live, lrow, j, v);
  line_fft<T, NX, +1, 1, 1>(v, j, lrow, twc);
  if (!live) return;
  if (react) {
    const T F = (T)react[inst].F, kk = (T)react[inst].k, hdt = (T)(0.5 * react[inst].dt);
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      T ru, rv;
      gs_react(v[m].re, v[m].im, F, kk, ru, rv);
      v[m] = mk(v[m].re + hdt * ru, v[m].im + hdt * rv);
    }
  }
  T* ra = a + inst * istride + (long long)y * pitch;
  T* rb = b ? b + inst * istride + (long long)y * pitch : nullptr;
  if (keep != (T)0) {
    const T fresh = (T)1 - keep;
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const int x = j + C::TN * m;
      v[m].re = keep * ra[x] + fresh * v[m].re;
      if (rb) v[m].im = keep * rb[x] + fresh * v[m].im;
    }
  }
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int x = j + C::TN * m;
    ra[x] = v[m].re;
    if (rb) rb[x] = v[m].im;
  }
  if (ghosts) {
    // Periodic ghost ring, x first then y (moljax/core/bc.py:202-215): row y
    // is also written to the ghost row on the opposite side when it is the
    // first/last interior row; the x ghosts follow the row.
    const long long offs[3] = {0, (y == 0) ? (long long)ny * pitch : 0,
                               (y == ny - 1) ? -(long long)ny * pitch : 0};
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      if (o > 0 && offs[o] == 0) continue;
      T* qa = ra + offs[o];
      T* qb = rb ? rb + offs[o] : nullptr;
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int x = j + C::TN * m;
        if (o > 0) {
          qa[x] = v[m].re;
          if (qb) qb[x] = v[m].im;
        }
        if (x == 0) {
          qa[NX] = v[m].re;
          if (qb) qb[NX] = v[m].im;
        }
        if (x == NX - 1) {
          qa[-1] = v[m].re;
          if (qb) qb[-1] = v[m].im;
        }
      }
    }
  }
}

// Fused Gray-Scott row pass: packed spectra -> physical (u, v) -> explicit
// reaction step -> packed spectra, in place.
// RXN = RXN_HEUN (IMEX-Strang), as moljax/core/stepping.py:414-425:
//   R1 = R(y*), y~ = y* + dt R1, R2 = R(y~), y** = y* + (dt/2)(R1 + R2).
// RXN = RXN_EULER (IMEX-Euler / SSPRK2 stages, stepping.py:351-360,472-489):
//   y + dt R(y); with rx = 1 y + (dt/2) R(y), with rx = 2 two such half
//   steps in a row (the reaction half-steps that close one step and open
//   the next of benchmark_gray_scott.py:268-289).
//
// A workgroup takes NT consecutive row tiles of one instance and issues the
// loads of all of them first, so that the later tiles' loads are in flight
// while the first is transformed (a 512-point line gives a thread only 8
// values, too few bytes in flight with one tile).  The arithmetic of every
// row is the same for any NT.
template <typename T, int NX, int TR, int RXN, int NT>
__global__ __launch_bounds__((RowCfg<NX, TR>::BT)) void k_row_gs(cpx<T>* __restrict__ W,
                                                                  const InstParams* __restrict__ ip, int ny,
                                                                  int groups_per_inst,
                                                                  const cpx<T>* __restrict__ tw, int mstep,
                                                                  int rx) {
  using C = RowCfg<NX, TR>;
  __shared__ cpx<T> lds[C::TR * C::LP];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / groups_per_inst;
  const int tile0 = (w - inst * groups_per_inst) * NT;
  // masked chain step: an instance past its own step count is left as is
  if (mstep >= 0 && mstep >= (int)ip[inst].aux) return;
  const int j = threadIdx.x % C::TN, rl = threadIdx.x / C::TN;
  cpx<T>* lrow = lds + rl * C::LP;
  const T F = (T)ip[inst].F, kk = (T)ip[inst].k, dt = (T)ip[inst].dt;
  const T hdt = (T)(0.5 * ip[inst].dt);
  TwCache<T, NX> twc;
  load_twiddles(twc, j, tw);
  // rows past ny (the last tile when TR does not divide ny): transforms only
  bool live[NT];
  cpx<T>* wrow[NT];
  cpx<T> g[NT > 1 ? NT : 1][C::P];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int y = (tile0 + t) * C::TR + rl;
    live[t] = y < ny;
    wrow[t] = W + ((long long)inst * ny + (live[t] ? y : 0)) * NX;
    if constexpr (NT > 1) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) g[t][m] = live[t] ? wrow[t][j + C::TN * m] : mk((T)0, (T)0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    cpx<T> v[C::P];
    if constexpr (NT == 1) {
      load_unpack<T, NX, TR>(wrow[0], live[0], lrow, j, v);
    } else {
      if (t > 0) __syncthreads();  // the previous tile's pack reads of LDS are done
#pragma unroll
      for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, g[t][m]);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < C::P; ++m) v[m] = unpack_at<T, NX>(lrow, j + C::TN * m);
      __syncthreads();
    }
    line_fft<T, NX, +1, 1, 1>(v, j, lrow, twc);
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const T u0 = v[m].re, v0 = v[m].im;
      T r1u, r1v, r2u, r2v;
      gs_react(u0, v0, F, kk, r1u, r1v);
      if constexpr (RXN == RXN_EULER) {
        if (rx == 0) {
          v[m] = mk(u0 + dt * r1u, v0 + dt * r1v);
        } else {
          T ut = u0 + hdt * r1u, vt = v0 + hdt * r1v;
          if (rx == 2) {
            gs_react(ut, vt, F, kk, r2u, r2v);
            ut = ut + hdt * r2u;
            vt = vt + hdt * r2v;
          }
          v[m] = mk(ut, vt);
        }
      } else {
        const T ut = u0 + dt * r1u, vt = v0 + dt * r1v;
        gs_react(ut, vt, F, kk, r2u, r2v);
        v[m] = mk(u0 + hdt * (r1u + r2u), v0 + hdt * (r1v + r2v));
      }
    }
    line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
#pragma unroll
    for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, v[m]);
    __syncthreads();
    if (live[t]) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int p = j + C::TN * m;
        wrow[t][p] = pack_at<T, NX>(lrow, p);
      }
    }
  }
}

// Increment-form IMEX-Strang row pass.  The state stays physical in the
// caller's arrays, held during the chain as (w, v) = (u - 1, v), its
// deviation from (1, 0), the homogeneous steady state of Gray-Scott; W
// carries the packed x-spectra of a diffusion increment:
//   INC_FIRST  w <- u - 1 (written back); W <- x-FFT of (w, v)
//   INC_MID    (w, v) += x-iFFT(W); Heun reaction step of R written for w
//              (u = w + 1); W <- x-FFT of (w, v)
//   INC_LAST   (w, v) += x-iFFT(W); u <- w + 1; optional periodic ghost ring
// Between row passes the column pass applies s (g - 1) or s (g^2 - 1)
// (NREP_INC1 / NREP_INC2), so (w, v) + x-iFFT(W) is H (w, v) or H^2 (w, v):
// the operator of mjx_gs_imex_strang (H leaves constants unchanged).  Only
// the increment passes through the transforms, so their rounding scales
// with it instead of with the state, and the stored deviation rounds
// relative to |u - 1| instead of |u|.
template <typename T, int NX, int TR, int MODE>
__global__ __launch_bounds__((RowCfg<NX, TR>::BT)) void k_row_gs_inc(cpx<T>* __restrict__ W, T* __restrict__ a,
                                                                      T* __restrict__ b, long long pitch,
                                                                      long long istride,
                                                                      const InstParams* __restrict__ ip, int ny,
                                                                      int rows_per_inst_tiles, int ghosts,
                                                                      const cpx<T>* __restrict__ tw) {
  using C = RowCfg<NX, TR>;
  __shared__ cpx<T> lds[C::TR * C::LP];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / rows_per_inst_tiles;
  const int tile = w - inst * rows_per_inst_tiles;
  const int j = threadIdx.x % C::TN, rl = threadIdx.x / C::TN;
  const int y = tile * C::TR + rl;
  const bool live = y < ny;  // rows past ny: transforms only
  cpx<T>* lrow = lds + rl * C::LP;
  cpx<T>* wrow = W + ((long long)inst * ny + (live ? y : 0)) * NX;
  T* ra = a + inst * istride + (long long)(live ? y : 0) * pitch;
  T* rb = b + inst * istride + (long long)(live ? y : 0) * pitch;
  TwCache<T, NX> twc;
  load_twiddles(twc, j, tw);
  T su[C::P], sv[C::P];  // (w, v)
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int x = j + C::TN * m;
    su[m] = live ? (MODE == INC_FIRST ? ra[x] - (T)1 : ra[x]) : (T)0;
    sv[m] = live ? rb[x] : (T)0;
  }
  if constexpr (MODE == INC_FIRST) {
    if (live) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) ra[j + C::TN * m] = su[m];
    }
  }
  cpx<T> v[C::P];
  if constexpr (MODE != INC_FIRST) {
    load_unpack<T, NX, TR>(wrow, live, lrow, j, v);
    line_fft<T, NX, +1, 1, 1>(v, j, lrow, twc);
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      su[m] += v[m].re;
      sv[m] += v[m].im;
    }
  }
  if constexpr (MODE == INC_MID) {
    const T F = (T)ip[inst].F, kk = (T)ip[inst].k, dt = (T)ip[inst].dt;
    const T hdt = (T)(0.5 * ip[inst].dt);
    // R of moljax/core/operators.py:655-667 with u = w + 1:
    // (-u v^2 - F w, u v^2 - (F + k) v)
    auto react_w = [&](T w_, T v_, T& rw, T& rv) {
      const T uvv = (w_ + (T)1) * v_ * v_;
      rw = -uvv - F * w_;
      rv = uvv - (F + kk) * v_;
    };
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      T r1u, r1v, r2u, r2v;
      react_w(su[m], sv[m], r1u, r1v);
      const T ut = su[m] + dt * r1u, vt = sv[m] + dt * r1v;
      react_w(ut, vt, r2u, r2v);
      su[m] = su[m] + hdt * (r1u + r2u);
      sv[m] = sv[m] + hdt * (r1v + r2v);
    }
  }
  if constexpr (MODE == INC_LAST) {
#pragma unroll
    for (int m = 0; m < C::P; ++m) su[m] += (T)1;
  }
  if constexpr (MODE != INC_FIRST) {
    if (live) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int x = j + C::TN * m;
        ra[x] = su[m];
        rb[x] = sv[m];
      }
    }
  }
  if constexpr (MODE == INC_LAST) {
    if (!live || !ghosts) return;
    // periodic ghost ring as k_row_inv_unpack writes it (x ghosts with the
    // row, first/last interior rows also into the opposite ghost row)
    const long long offs[3] = {0, (y == 0) ? (long long)ny * pitch : 0, (y == ny - 1) ? -(long long)ny * pitch : 0};
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      if (o > 0 && offs[o] == 0) continue;
      T* qa = ra + offs[o];
      T* qb = rb + offs[o];
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int x = j + C::TN * m;
        if (o > 0) {
          qa[x] = su[m];
          qb[x] = sv[m];
        }
        if (x == 0) {
          qa[NX] = su[m];
          qb[NX] = sv[m];
        }
        if (x == NX - 1) {
          qa[-1] = su[m];
          qb[-1] = sv[m];
        }
      }
    }
    return;
  } else {
#pragma unroll
    for (int m = 0; m < C::P; ++m) v[m] = mk(su[m], sv[m]);
    line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
#pragma unroll
    for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, v[m]);
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const int p = j + C::TN * m;
      wrow[p] = pack_at<T, NX>(lrow, p);
    }
  }
}

// ----------------------------------------------------------- column pass
template <int NY, int TC_>
struct ColCfg {
  static constexpr int P = plan_p(NY);
  static constexpr int TN = NY / P;
  static constexpr int TC = TC_;      // columns per workgroup
  static constexpr int BT = TC * TN;  // threads per workgroup
};

template <typename T, int NY, int TC_, bool LAM2D, int NREP>
__global__ __launch_bounds__((ColCfg<NY, TC_>::BT)) void k_col_apply(
    cpx<T>* __restrict__ W, int nx, const T* __restrict__ lamx, const T* __restrict__ lamy,
    const T* __restrict__ lam2d, const InstParams* __restrict__ ip, int nrep_arg, T s, int col_tiles,
    const cpx<T>* __restrict__ tw, int mstep) {
  using C = ColCfg<NY, TC_>;
  constexpr int TC = C::TC;
  __shared__ cpx<T> lds[NY * TC];
  __shared__ T lds_lamy[LAM2D ? 1 : NY];
  const int w = xcd_remap_rev(blockIdx.x, gridDim.x);
  const int inst = w / col_tiles;
  const int tile = w - inst * col_tiles;
  // Per-instance step counts (instance row aux): NREP_PER_INSTANCE raises the
  // multiplier to that power; a masked chain step mstep >= 0 applies H^2
  // before the instance's last step, H at it, and leaves the instance
  // untouched after it; MSTEP_ACTIVE leaves instances with aux <= 0
  // untouched (the whole workgroup belongs to one instance).
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
  const int c = threadIdx.x % TC, j = threadIdx.x / TC;
  const int p0 = tile * TC;
  const int p = p0 + c;
  const int half = nx >> 1;
  // the last tile may hold columns past nx (nx/2 not a multiple of TC): they
  // run the transforms on zeros, for the barriers, and store nothing
  const bool live = p < nx;
  cpx<T>* col = W + (long long)inst * NY * nx + (live ? p : 0);
  cpx<T> v[C::P];
#pragma unroll
  for (int m = 0; m < C::P; ++m) v[m] = live ? col[(long long)(j + C::TN * m) * nx] : mk((T)0, (T)0);
  // The exponential multiplier e^{beta (lx + ly)} = e^{beta lx} e^{beta ly}
  // stages e^{beta ly} in LDS (NY exponentials per workgroup, not per thread)
  // for the field of the tile's first column; when nx/2 is a multiple of TC
  // every column of a tile belongs to that field, otherwise the columns of
  // the other field in the one tile that straddles nx/2 take e^{beta ly}
  // directly.
  const bool expmode = NREP == 0 && !LAM2D && nrep == -2;
  const bool f0 = p0 >= half;
  if constexpr (!LAM2D) {
    const T beta_w = (T)(f0 ? ip[inst].beta_b : ip[inst].beta_a);
    for (int i = threadIdx.x; i < NY; i += C::BT) lds_lamy[i] = expmode ? exp(beta_w * lamy[i]) : lamy[i];
  }
  TwCache<T, NY> twc;
  load_twiddles(twc, j, tw);
  line_fft<T, NY, -1, TC, 0>(v, j, lds + c, twc);  // its barriers also publish lds_lamy

  const bool fb = p >= half;
  const int kx = !live ? 0 : fb ? p - half : p;
  const T alpha = (T)(fb ? ip[inst].alpha_b : ip[inst].alpha_a);
  const T beta = (T)(fb ? ip[inst].beta_b : ip[inst].beta_a);
  const bool staged = fb == f0;
  auto ey = [&](int ky) -> T { return staged ? lds_lamy[ky] : exp(beta * lamy[ky]); };
  auto multiply_plain = [&]() {
    if (!live) return;
    const T lx = LAM2D ? (T)0 : lamx[kx];
    if (expmode) {
      const T ex = exp(beta * lx) * s;
#pragma unroll
      for (int m = 0; m < C::P; ++m) v[m] = scale(v[m], ex * ey(j + C::TN * m));
      return;
    }
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const int ky = j + C::TN * m;
      T lam;
      if constexpr (LAM2D) lam = lam2d[(long long)ky * nx + kx];
      else lam = lx + lds_lamy[ky];
      v[m] = scale(v[m], spec_mult<T, NREP>(lam, alpha, beta, nrep, s));
    }
  };
  if (p0 == 0 || (p0 <= half && half < p0 + TC)) {
    // This tile holds packed column 0 or nx/2: two real sequences with the
    // multipliers of kx = 0 (ma) and kx = nx/2 (mb),
    // out(ky) = (ma+mb)/2 Z(ky) + (ma-mb)/2 conj(Z(-ky)).
#pragma unroll
    for (int m = 0; m < C::P; ++m) lds[(j + C::TN * m) * TC + c] = v[m];
    __syncthreads();
    if (p == 0 || p == half) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int ky = j + C::TN * m;
        const int kym = ky == 0 ? 0 : NY - ky;
        const cpx<T> zm = conjg(lds[kym * TC + c]);
        T la, lb, ma, mb;
        if constexpr (LAM2D) {
          la = lam2d[(long long)ky * nx + 0];
          lb = lam2d[(long long)ky * nx + half];
        } else {
          la = lamx[0] + lds_lamy[ky];
          lb = lamx[half] + lds_lamy[ky];
        }
        if (expmode) {
          ma = exp(beta * lamx[0]) * s * ey(ky);
          mb = exp(beta * lamx[half]) * s * ey(ky);
        } else {
          ma = spec_mult<T, NREP>(la, alpha, beta, nrep, s);
          mb = spec_mult<T, NREP>(lb, alpha, beta, nrep, s);
        }
        const T cp = (T)0.5 * (ma + mb), cm = (T)0.5 * (ma - mb);
        v[m] = mk(cp * v[m].re + cm * zm.re, cp * v[m].im + cm * zm.im);
      }
    } else {
      multiply_plain();
    }
    __syncthreads();
  } else {
    multiply_plain();
  }
  line_fft<T, NY, +1, TC, 0>(v, j, lds + c, twc);
  if (!live) return;
#pragma unroll
  for (int m = 0; m < C::P; ++m) col[(long long)(j + C::TN * m) * nx] = v[m];
}

// ------------------------------------------------------------- dispatch
// Lines per workgroup: about 256 threads in FP32 and 128 in FP64 (LDS), at
// least one full line, and never more lines than the grid has (rows: ny,
// column tile: nx/2, so that with power-of-two sides the packed special
// columns 0 and nx/2 fall at tile starts).  Mixed-radix sides may leave a
// partial last tile; the kernels skip its missing lines.
template <typename T, int N>
constexpr int lines_default() {
  constexpr int bt = sizeof(T) == 8 ? 128 : 256;
  return (N / plan_p(N)) >= bt ? 1 : bt / (N / plan_p(N));
}

// The largest power of two no larger than def or limit (with_lines
// instantiates powers of two).
inline int pick_lines(int def, int limit) {
  const int l = def < limit ? def : limit;
  int p = 1;
  while (2 * p <= l) p *= 2;
  return p;
}

// Small batches: halve the lines per workgroup until a launch has about 512
// workgroups (two per CU), so that one instance still spreads over the chip.
// Each line's arithmetic does not depend on the tile, so results are the same.
inline int fill_lines(int l, long long lines_total) {
  while (l > 1 && lines_total / l < 512) l >>= 1;
  return l;
}

// Tuned lines per workgroup for a kernel family, else the default; always a
// power of two no larger than the default (what with_lines instantiates).
template <typename T, int N>
inline int lines_for(int kind) {
  const int t = tuned_lines(kind, sizeof(T) == 8 ? MJX_F64 : MJX_F32, N);
  const int d = lines_default<T, N>();
  return (t > 0 && t <= d) ? t : d;
}

template <typename T, int N, int L>
struct RowLaunch {
  static int fwd_pack(int B, int ny, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                      T in_scale, hipStream_t st) {
    const int tiles = (ny + L - 1) / L;
    hipLaunchKernelGGL((k_row_fwd_pack<T, N, L>), dim3(tiles * B), dim3(RowCfg<N, L>::BT), 0, st, a, b, pitch,
                       istride, W, ny, tiles, in_scale, twiddles<T>() + tw_base(N));
    return check_launch("row_fwd_pack");
  }
  static int inv_unpack(int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                        int ghosts, T keep, const InstParams* react, hipStream_t st) {
    const int tiles = (ny + L - 1) / L;
    hipLaunchKernelGGL((k_row_inv_unpack<T, N, L>), dim3(tiles * B), dim3(RowCfg<N, L>::BT), 0, st, W, a, b,
                       pitch, istride, ny, tiles, ghosts, keep, twiddles<T>() + tw_base(N), react);
    return check_launch("row_inv_unpack");
  }
  static int gs(int B, int ny, cpx<T>* W, const InstParams* ip, int rxn, int mstep, hipStream_t st) {
    const int tiles = (ny + L - 1) / L;
    // Two row tiles per workgroup (the second tile's loads in flight during
    // the first) when the grid keeps >= 2048 workgroups: FP32 lines of at
    // most 8 values a thread (n >= 512), and 256-point lines once the packed
    // plane outgrows the 256 MiB Infinity Cache, where the row pass reads
    // from HBM (256^2: B = 2048 gains 4.5 %; B = 512, whose plane the column
    // pass leaves in that cache, loses 1.5 %).
    constexpr bool two_ok = sizeof(T) == 4 && (N == 256 || (N >= 512 && plan_p(N) <= 8));
    if constexpr (two_ok) {
      const bool two_here = N != 256 || (long long)B * ny * N * (long long)sizeof(cpx<T>) > (256LL << 20);
      if (two_here && tiles % 2 == 0 && (long long)(tiles / 2) * B >= 2048) {
        if (rxn != RXN_HEUN)
          hipLaunchKernelGGL((k_row_gs<T, N, L, RXN_EULER, 2>), dim3(tiles / 2 * B), dim3(RowCfg<N, L>::BT), 0, st,
                             W, ip, ny, tiles / 2, twiddles<T>() + tw_base(N), mstep, rxn - RXN_EULER);
        else
          hipLaunchKernelGGL((k_row_gs<T, N, L, RXN_HEUN, 2>), dim3(tiles / 2 * B), dim3(RowCfg<N, L>::BT), 0, st,
                             W, ip, ny, tiles / 2, twiddles<T>() + tw_base(N), mstep, 0);
        return check_launch("row_gs");
      }
    }
    if (rxn != RXN_HEUN)
      hipLaunchKernelGGL((k_row_gs<T, N, L, RXN_EULER, 1>), dim3(tiles * B), dim3(RowCfg<N, L>::BT), 0, st, W, ip,
                         ny, tiles, twiddles<T>() + tw_base(N), mstep, rxn - RXN_EULER);
    else
      hipLaunchKernelGGL((k_row_gs<T, N, L, RXN_HEUN, 1>), dim3(tiles * B), dim3(RowCfg<N, L>::BT), 0, st, W, ip,
                         ny, tiles, twiddles<T>() + tw_base(N), mstep, 0);
    return check_launch("row_gs");
  }
  static int gs_inc(int B, int ny, cpx<T>* W, T* a, T* b, long long pitch, long long istride, const InstParams* ip,
                    int mode, int ghosts, hipStream_t st) {
    const int tiles = (ny + L - 1) / L;
    const dim3 g(tiles * B), bt(RowCfg<N, L>::BT);
    const cpx<T>* tw = twiddles<T>() + tw_base(N);
    if (mode == INC_FIRST)
      hipLaunchKernelGGL((k_row_gs_inc<T, N, L, INC_FIRST>), g, bt, 0, st, W, a, b, pitch, istride, ip, ny, tiles,
                         ghosts, tw);
    else if (mode == INC_MID)
      hipLaunchKernelGGL((k_row_gs_inc<T, N, L, INC_MID>), g, bt, 0, st, W, a, b, pitch, istride, ip, ny, tiles,
                         ghosts, tw);
    else
      hipLaunchKernelGGL((k_row_gs_inc<T, N, L, INC_LAST>), g, bt, 0, st, W, a, b, pitch, istride, ip, ny, tiles,
                         ghosts, tw);
    return check_launch("row_gs_inc");
  }
};

template <typename T, int N, int L>
struct ColLaunch {
  template <bool LAM2D, int NREP>
  static void go(int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d, const InstParams* ip,
                 int nrep, T s, int mstep, hipStream_t st) {
    const int tiles = (nx + L - 1) / L;
    hipLaunchKernelGGL((k_col_apply<T, N, L, LAM2D, NREP>), dim3(tiles * B), dim3(ColCfg<N, L>::BT), 0, st, W,
                       nx, lamx, lamy, lam2d, ip, nrep, s, tiles, twiddles<T>() + tw_base(N), mstep);
  }
  // compile-time increment forms of the increment-form Strang chain, which
  // runs on power-of-two lengths only
  template <int NREP>
  static void go_pow2(int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d, const InstParams* ip,
                      int nrep, T s, int mstep, hipStream_t st) {
    if constexpr (pow2_len(N)) go<false, NREP>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
    else go<false, 0>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
  }
  static int apply(int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d, const InstParams* ip,
                   int nrep, T s, int mstep, hipStream_t st) {
    // per-instance powers (masked steps, NREP_PER_INSTANCE) use the runtime power path
    const int fixed = mstep >= 0 ? 0 : nrep;
    if (lam2d) {
      if (fixed == 1) go<true, 1>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else if (fixed == 2) go<true, 2>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else go<true, 0>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
    } else {
      if (fixed == 1) go<false, 1>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else if (fixed == 2) go<false, 2>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else if (fixed == -1) go<false, -1>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else if (pow2_len(N) && fixed == NREP_INC1) go_pow2<NREP_INC1>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else if (pow2_len(N) && fixed == NREP_INC2) go_pow2<NREP_INC2>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      else go<false, 0>(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
    }
    return check_launch("col_apply");
  }
};

// Line counts instantiated for a length: every power of two up to the
// default for the power-of-two lengths (the autotuner picks among them); for
// the mixed-radix lengths only the default rounded down to a power of two and
// its quarters, sixteenths, ... (enough for small batches, at a third of the
// code).
__host__ __device__ constexpr int pow2_floor(int d) {
  int p = 1;
  while (2 * p <= d) p *= 2;
  return p;
}
template <typename T, int N>
constexpr bool lines_built(int l) {
  constexpr int D = lines_default<T, N>();
  if (l > D) return false;
  if (!mixed_len(N) || l == 1) return true;
  int q = pow2_floor(D);
  while (q > l) q /= 4;
  return q == l;
}

// Calls F<L>::run(args...) for the runtime line count l (a power of two no
// larger than lines_default<T, N>()), rounded down to a built count.
template <typename T, int N, template <int> class F, typename... A>
int with_lines(int l, A... args) {
  while (l > 1 && !lines_built<T, N>(l)) l /= 2;
  if constexpr (lines_built<T, N>(64)) if (l == 64) return F<64>::run(args...);
  if constexpr (lines_built<T, N>(32)) if (l == 32) return F<32>::run(args...);
  if constexpr (lines_built<T, N>(16)) if (l == 16) return F<16>::run(args...);
  if constexpr (lines_built<T, N>(8)) if (l == 8) return F<8>::run(args...);
  if constexpr (lines_built<T, N>(4)) if (l == 4) return F<4>::run(args...);
  if constexpr (lines_built<T, N>(2)) if (l == 2) return F<2>::run(args...);
  if (l == 1) return F<1>::run(args...);
  return fail(MJX_E_SIZE, "no kernel for tile of " + std::to_string(l) + " lines");
}

#define MJX_FOR_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) MJX_FOR_MIXED(X)

// The launchers of one precision are compiled in three translation units,
// one per set of line lengths, so that they build in parallel.
enum { SET_POW2 = 0, SET_MIX5 = 1, SET_MIX3 = 2 };
__host__ __device__ constexpr int size_set(int n) {
  return pow2_len(n) ? SET_POW2 : (n % 5 == 0) ? SET_MIX5 : SET_MIX3;
}

template <typename T, int SET>
struct Launch {
  template <int N>
  struct FwdPack {
    template <int L>
    struct F {
      static int run(int B, int ny, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                     T in_scale, hipStream_t st) {
        return RowLaunch<T, N, L>::fwd_pack(B, ny, a, b, pitch, istride, W, in_scale, st);
      }
    };
  };
  template <int N>
  struct InvUnpack {
    template <int L>
    struct F {
      static int run(int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch, long long istride, int ghosts,
                     T keep, const InstParams* react, hipStream_t st) {
        return RowLaunch<T, N, L>::inv_unpack(B, ny, W, a, b, pitch, istride, ghosts, keep, react, st);
      }
    };
  };
  template <int N>
  struct Gs {
    template <int L>
    struct F {
      static int run(int B, int ny, cpx<T>* W, const InstParams* ip, int rxn, int mstep, hipStream_t st) {
        return RowLaunch<T, N, L>::gs(B, ny, W, ip, rxn, mstep, st);
      }
    };
  };
  template <int N>
  struct GsInc {
    template <int L>
    struct F {
      static int run(int B, int ny, cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                     const InstParams* ip, int mode, int ghosts, hipStream_t st) {
        return RowLaunch<T, N, L>::gs_inc(B, ny, W, a, b, pitch, istride, ip, mode, ghosts, st);
      }
    };
  };
  template <int N>
  struct Col {
    template <int L>
    struct F {
      static int run(int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                     const InstParams* ip, int nrep, T s, int mstep, hipStream_t st) {
        return ColLaunch<T, N, L>::apply(B, nx, W, lamx, lamy, lam2d, ip, nrep, s, mstep, st);
      }
    };
  };

  static int row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,
                          cpx<T>* W, hipStream_t st, T in_scale = (T)1) {
    switch (nx) {
#define X(N)                                 \
  case N:                                    \
    if constexpr (size_set(N) == SET)        \
    return with_lines<T, N, FwdPack<N>::template F>(pick_lines(fill_lines(lines_for<T, N>(TK_ROW), (long long)B * ny), ny), B, ny, a, b, pitch, istride, W, in_scale, st); \
    else break;
      MJX_FOR_SIZES(X)
#undef X
      default:
        return fail(MJX_E_SIZE, "unsupported nx " + std::to_string(nx));
    }
    return fail(MJX_E_SIZE, "nx " + std::to_string(nx) + " belongs to another launcher set");
  }
  static int row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,
                            long long istride, int ghosts, hipStream_t st, T keep = (T)0,
                            const InstParams* react = nullptr) {
    switch (nx) {
#define X(N)                                                                                              \
  case N:                                                                                                 \
    if constexpr (size_set(N) == SET)                                                                     \
    return with_lines<T, N, InvUnpack<N>::template F>(pick_lines(fill_lines(lines_for<T, N>(TK_ROW), (long long)B * ny), ny), B, ny, W, a, b, pitch, \
                                                   istride, ghosts, keep, react, st); \
    else break;
      MJX_FOR_SIZES(X)
#undef X
      default:
        return fail(MJX_E_SIZE, "unsupported nx " + std::to_string(nx));
    }
    return fail(MJX_E_SIZE, "nx " + std::to_string(nx) + " belongs to another launcher set");
  }
  static int row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn = RXN_HEUN,
                    int mstep = -1) {
    switch (nx) {
#define X(N)                          \
  case N:                             \
    if constexpr (size_set(N) == SET) \
    return with_lines<T, N, Gs<N>::template F>(pick_lines(fill_lines(lines_for<T, N>(TK_ROW), (long long)B * ny), ny), B, ny, W, ip, rxn, mstep, st); \
    else break;
      MJX_FOR_SIZES(X)
#undef X
      default:
        return fail(MJX_E_SIZE, "unsupported nx " + std::to_string(nx));
    }
    return fail(MJX_E_SIZE, "nx " + std::to_string(nx) + " belongs to another launcher set");
  }
  // The increment-form row pass is built for the power-of-two lengths only.
  static int row_gs_inc(int nx, int B, int ny, cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                        const InstParams* ip, int mode, int ghosts, hipStream_t st) {
    if constexpr (SET != SET_POW2) {
      return fail(MJX_E_SIZE, "increment-form IMEX-Strang: nx must be a power of two in [16, 4096]");
    } else {
      switch (nx) {
#define X(N)                                                                                                   \
  case N:                                                                                                      \
    return with_lines<T, N, GsInc<N>::template F>(pick_lines(fill_lines(lines_for<T, N>(TK_ROW), (long long)B * ny), ny), \
                                                  B, ny, W, a, b, pitch, istride, ip, mode, ghosts, st);
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
        default:
          return fail(MJX_E_SIZE, "increment-form IMEX-Strang: nx must be a power of two in [16, 4096]");
      }
    }
  }
  static int col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                       const InstParams* ip, int nrep, hipStream_t st, int mstep = -1) {
    const T s = (T)(1.0 / ((double)nx * (double)ny));
    switch (ny) {
#define X(N)                                                                                                 \
  case N:                                                                                                    \
    if constexpr (size_set(N) == SET)                                                                        \
    return with_lines<T, N, Col<N>::template F>(pick_lines(fill_lines(lines_for<T, N>(TK_COL), (long long)B * nx), nx / 2), B, nx, W, lamx, lamy, lam2d, \
                                             ip, nrep, s, mstep, st); \
    else break;
      MJX_FOR_SIZES(X)
#undef X
      default:
        return fail(MJX_E_SIZE, "unsupported ny " + std::to_string(ny));
    }
    return fail(MJX_E_SIZE, "ny " + std::to_string(ny) + " belongs to another launcher set");
  }
};

}  // namespace mjx
