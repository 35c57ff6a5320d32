// Any-length line transforms by the chirp-z (Bluestein) algorithm, for grid
// sizes the power-of-two Stockham kernels do not cover.
//
// One line of N points is transformed by one workgroup.  Every transform here
// is a generalised DFT with half-integer offsets,
//   y_k = sum_{n<N} x_n exp(-2 pi i (k + ko)(n + no) / P),   k < N,
// which with 2 K M = K^2 + M^2 - (K - M)^2 (K = k + ko, M = n + no) becomes
//   y_k = c_k * sum_n (x_n d_n) e_{k-n},
//   c_k = exp(-i pi (k+ko)^2 / P),  d_n = exp(-i pi (n+no)^2 / P),
//   e_m = exp(+i pi (m + ko - no)^2 / P),  |m| < N,
// a linear convolution of length 2N-1, done as a circular one of length
// L = 2^lg >= 2N - 1 with two radix-2 FFTs in LDS: decimation in time from
// bit-reversed order (forward), a pointwise product with FFT_L(e)/L, and
// decimation in frequency back to bit-reversed order (inverse, unscaled).
// The tables c, d and FFT_L(e)/L are evaluated in FP64 on the host (exact
// integer reduction of the squared offsets modulo 8P) and rounded once.
//
// The real transforms of the non-periodic solves, with N points per line:
//   DST-I   X_k = sum x_n sin(pi (k+1)(n+1)/(N+1))  P = 2(N+1), ko = no = 1, -Im
//   DCT-II  X_k = sum x_n cos(pi k (2n+1)/(2N))      P = 2N, ko = 0, no = 1/2, Re
//   DCT-III x_n = sum X_k cos(pi k (2n+1)/(2N))      P = 2N, ko = 1/2, no = 0, Re
//   DST-II  X_k = sum x_n sin(pi (k+1)(2n+1)/(2N))   P = 2N, ko = 1, no = 1/2, -Im
//   DST-III x_n = sum X_k sin(pi (k+1)(2n+1)/(2N))   P = 2N, ko = 1/2, no = 1, -Im
// (output index k carries ko, input index n carries no).  The periodic
// complex DFT is P = N, ko = no = 0; its inverse is conj(DFT(conj(x))).
#pragma once
#include <hip/hip_runtime.h>

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
  int n, L, lg;
};

// Largest number of line points one thread keeps in registers between the two
// transforms of a fused pass (n <= 8 * blockDim by the launch rule).
constexpr int CZT_KEEP = 8;

__device__ __forceinline__ int czt_brev(int i, int lg) {
  return lg ? (int)(__brev((unsigned)i) >> (32 - lg)) : 0;
}

// lds[brev(n)] <- x_n d_n for n < N and 0 for N <= n < L; x from get(n).
template <typename T, class G>
__device__ __forceinline__ void czt_load(cpx<T>* lds, const CztPlan<T>& p, G get) {
  for (int n = threadIdx.x; n < p.L; n += blockDim.x)
    lds[czt_brev(n, p.lg)] = n < p.n ? cmul(get(n), p.d[n]) : mk((T)0, (T)0);
}

// The circular convolution with e, in place; starts and ends with a barrier.
template <typename T>
__device__ void czt_conv(cpx<T>* lds, const CztPlan<T>& p, const cpx<T>* __restrict__ tw) {
  const int L = p.L, H = p.L >> 1;
  for (int half = 1; half < L; half <<= 1) {
    __syncthreads();
    const int ts = MJX_TWN / (2 * half);
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
      const int pos = j & (half - 1);
      const int i0 = ((j - pos) << 1) + pos, i1 = i0 + half;
      const cpx<T> t = cmul(tw[pos * ts], lds[i1]);
      const cpx<T> u = lds[i0];
      lds[i0] = u + t;
      lds[i1] = u - t;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < L; k += blockDim.x) lds[k] = cmul(lds[k], p.e[k]);
  for (int half = H; half >= 1; half >>= 1) {
    __syncthreads();
    const int ts = MJX_TWN / (2 * half);
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
      const int pos = j & (half - 1);
      const int i0 = ((j - pos) << 1) + pos, i1 = i0 + half;
      const cpx<T> u = lds[i0], v = lds[i1];
      lds[i0] = u + v;
      lds[i1] = cmul(u - v, conjg(tw[pos * ts]));
    }
  }
  __syncthreads();
}

// Output k (< N) of the transform whose convolution is in lds.
template <typename T>
__device__ __forceinline__ cpx<T> czt_out(const cpx<T>* lds, const CztPlan<T>& p, int k) {
  return cmul(p.c[k], lds[czt_brev(k, p.lg)]);
}

// Second transform of a fused pass from the values each thread kept for
// n = threadIdx.x + i * blockDim.x; starts with a barrier (the first
// transform's outputs have all been read).
template <typename T>
__device__ __forceinline__ void czt_reload(cpx<T>* lds, const CztPlan<T>& p, const cpx<T> (&keep)[CZT_KEEP]) {
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CZT_KEEP; ++i) {
    const int n = threadIdx.x + i * blockDim.x;
    if (n < p.n) lds[czt_brev(n, p.lg)] = cmul(keep[i], p.d[n]);
  }
  for (int n = p.n + threadIdx.x; n < p.L; n += blockDim.x) lds[czt_brev(n, p.lg)] = mk((T)0, (T)0);
}

extern __shared__ __align__(16) unsigned char czt_smem[];

// ------------------------------------------------------ real line transforms
enum { CZT_RE = 0, CZT_NEG_IM = 1 };

struct CztR2RArgs {
  const void* in;
  long long in_es, in_ls, in_is;
  void* out;
  long long out_es, out_ls, out_is;
  int nlines;
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

// One workgroup per line.  APPLY: forward transform, spectral multiplier,
// inverse transform (second plan); otherwise one transform times sc1.
template <typename T, bool APPLY>
__global__ __launch_bounds__(256) void k_czt_r2r(CztR2RArgs a, CztPlan<T> p1, CztPlan<T> p2,
                                                 const cpx<T>* __restrict__ tw) {
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(czt_smem);
  const int line = blockIdx.x % a.nlines, inst = blockIdx.x / a.nlines;
  const int N = p1.n;
  const T* in = (const T*)a.in + inst * a.in_is + (long long)line * a.in_ls;
  T* out = (T*)a.out + inst * a.out_is + (long long)line * a.out_ls;
  double lw = 0.0;
  if (a.lift_on) lw = a.mmode == R2R_POISSON ? -1.0 : a.inst[8 * inst];
  czt_load(lds, p1, [&](int n) {
    T x = in[(long long)n * a.in_es];
    if (a.lift_on) {
      const double add = (n == 0 ? a.lift[0] : 0.0) + (n == N - 1 ? a.lift[1] : 0.0) +
                         (line == 0 ? a.lift[2] : 0.0) + (line == a.nlines - 1 ? a.lift[3] : 0.0);
      x += (T)(lw * add);
    }
    return mk(x, (T)0);
  });
  czt_conv(lds, p1, tw);
  if constexpr (!APPLY) {
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const cpx<T> y = czt_out(lds, p1, k);
      const T v = (a.part1 == CZT_RE ? y.re : -y.im) * (T)a.sc1;
      T* o = out + (long long)k * a.out_es;
      *o = a.accumulate ? *o + v : v;
    }
  } else {
    const double* ip = a.inst ? a.inst + 8 * inst : nullptr;
    const T beta = ip ? (T)ip[0] : (T)a.beta0, alpha = ip ? (T)ip[2] : (T)0;
    const int power = (ip && a.mmode == R2R_HELM_POW) ? (int)ip[7] : 1;
    const T lc = a.lam_cross ? ((const T*)a.lam_cross)[line] : (T)0;
    const double fc = a.n_cross > 0 ? czt_norm(a.kind, line, a.n_cross) : 1.0;
    cpx<T> keep[CZT_KEEP];
#pragma unroll
    for (int i = 0; i < CZT_KEEP; ++i) {
      const int k = threadIdx.x + i * blockDim.x;
      keep[i] = mk((T)0, (T)0);
      if (k < N) {
        const cpx<T> y = czt_out(lds, p1, k);
        const T v = a.part1 == CZT_RE ? y.re : -y.im;
        const T m = r2r_mult<T>(((const T*)a.lam_line)[k] + lc, alpha, beta, a.mmode, power) *
                    (T)(czt_norm(a.kind, k, N) * fc);
        keep[i] = mk(v * m, (T)0);
      }
    }
    czt_reload(lds, p2, keep);
    czt_conv(lds, p2, tw);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const cpx<T> y = czt_out(lds, p2, k);
      const T v = (a.part2 == CZT_RE ? y.re : -y.im) * (T)a.sc2;
      T* o = out + (long long)k * a.out_es;
      *o = a.accumulate ? *o + v : v;
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

// Half-spectra of the two real lines packed in Z (the transform in lds).
template <typename T>
__device__ __forceinline__ void czt_split_store(const cpx<T>* lds, const CztPlan<T>& p, cpx<T>* __restrict__ wrow,
                                                int nx, int nh) {
  for (int k = threadIdx.x; k < nh; k += blockDim.x) {
    const int km = k == 0 ? 0 : nx - k;
    const cpx<T> z = czt_out(lds, p, k), zm = conjg(czt_out(lds, p, km));
    wrow[k] = mk((T)0.5 * (z.re + zm.re), (T)0.5 * (z.im + zm.im));
    wrow[nh + k] = mk((T)0.5 * (z.im - zm.im), (T)0.5 * (zm.re - z.re));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_czt_row_fwd_pack(const T* __restrict__ a, const T* __restrict__ b,
                                                          long long pitch, long long istride, cpx<T>* __restrict__ W,
                                                          int ny, int nh, T in_scale, CztPlan<T> px,
                                                          const cpx<T>* __restrict__ tw) {
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(czt_smem);
  const int y = blockIdx.x % ny, inst = blockIdx.x / ny;
  const T* ra = a + inst * istride + (long long)y * pitch;
  const T* rb = b ? b + inst * istride + (long long)y * pitch : nullptr;
  czt_load(lds, px, [&](int n) { return mk(ra[n] * in_scale, rb ? rb[n] * in_scale : (T)0); });
  czt_conv(lds, px, tw);
  czt_split_store(lds, px, W + ((long long)inst * ny + y) * 2 * nh, px.n, nh);
}

template <typename T>
__global__ __launch_bounds__(256) void k_czt_row_inv_unpack(const cpx<T>* __restrict__ W, T* __restrict__ a,
                                                            T* __restrict__ b, long long pitch, long long istride,
                                                            int ny, int nh, int ghosts, T keep_w,
                                                            const InstParams* __restrict__ react, CztPlan<T> px,
                                                            const cpx<T>* __restrict__ tw) {
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(czt_smem);
  const int y = blockIdx.x % ny, inst = blockIdx.x / ny;
  const int nx = px.n;
  const cpx<T>* wrow = W + ((long long)inst * ny + y) * 2 * nh;
  czt_load(lds, px, [&](int k) { return conjg(czt_merge(wrow, k, nx, nh)); });
  czt_conv(lds, px, tw);
  T* ra = a + inst * istride + (long long)y * pitch;
  T* rb = b ? b + inst * istride + (long long)y * pitch : nullptr;
  const long long offs[3] = {0, (y == 0) ? (long long)ny * pitch : 0, (y == ny - 1) ? -(long long)ny * pitch : 0};
  for (int x = threadIdx.x; x < nx; x += blockDim.x) {
    cpx<T> z = conjg(czt_out(lds, px, x));
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

// Fused Gray-Scott row pass on the half-spectra, in place: inverse x-DFT,
// explicit reaction (RXN_HEUN, or forward Euler over dt, dt/2 or two dt/2
// steps), forward x-DFT.
template <typename T>
__global__ __launch_bounds__(256) void k_czt_row_gs(cpx<T>* __restrict__ W, const InstParams* __restrict__ ip,
                                                    int ny, int nh, int rxn, int mstep, CztPlan<T> px,
                                                    const cpx<T>* __restrict__ tw) {
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(czt_smem);
  const int y = blockIdx.x % ny, inst = blockIdx.x / ny;
  if (mstep >= 0 && mstep >= (int)ip[inst].aux) return;
  const int nx = px.n;
  cpx<T>* wrow = W + ((long long)inst * ny + y) * 2 * nh;
  const T F = (T)ip[inst].F, kk = (T)ip[inst].k, dt = (T)ip[inst].dt, hdt = (T)(0.5 * ip[inst].dt);
  czt_load(lds, px, [&](int k) { return conjg(czt_merge(wrow, k, nx, nh)); });
  czt_conv(lds, px, tw);
  cpx<T> keep[CZT_KEEP];
#pragma unroll
  for (int i = 0; i < CZT_KEEP; ++i) {
    const int x = threadIdx.x + i * blockDim.x;
    keep[i] = mk((T)0, (T)0);
    if (x < nx) {
      const cpx<T> z = conjg(czt_out(lds, px, x));
      const T u0 = z.re, v0 = z.im;
      T r1u, r1v, r2u, r2v;
      gs_react(u0, v0, F, kk, r1u, r1v);
      cpx<T> o;
      if (rxn == RXN_HEUN) {
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
      keep[i] = o;
    }
  }
  czt_reload(lds, px, keep);
  czt_conv(lds, px, tw);
  czt_split_store(lds, px, wrow, nx, nh);
}

// Column pass over the 2*nh half-spectrum columns: y-DFT, multiplier
// spec_mult(lam, alpha_f, beta_f, nrep) * s, inverse y-DFT, in place.  The
// per-instance step rules are those of the power-of-two column pass.
template <typename T>
__global__ __launch_bounds__(256) void k_czt_col_apply(cpx<T>* __restrict__ W, int nx, int nh,
                                                       const T* __restrict__ lamx, const T* __restrict__ lamy,
                                                       const T* __restrict__ lam2d,
                                                       const InstParams* __restrict__ ip, int nrep_arg, T s,
                                                       int mstep, CztPlan<T> py, const cpx<T>* __restrict__ tw) {
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(czt_smem);
  const int cols = 2 * nh;
  const int c = blockIdx.x % cols, inst = blockIdx.x / cols;
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
  const int ny = py.n;
  const bool fb = c >= nh;
  const int kx = fb ? c - nh : c;
  const T alpha = (T)(fb ? ip[inst].alpha_b : ip[inst].alpha_a);
  const T beta = (T)(fb ? ip[inst].beta_b : ip[inst].beta_a);
  cpx<T>* col = W + (long long)inst * ny * cols + c;
  czt_load(lds, py, [&](int ky) { return col[(long long)ky * cols]; });
  czt_conv(lds, py, tw);
  cpx<T> keep[CZT_KEEP];
#pragma unroll
  for (int i = 0; i < CZT_KEEP; ++i) {
    const int ky = threadIdx.x + i * blockDim.x;
    keep[i] = mk((T)0, (T)0);
    if (ky < ny) {
      const T lam = lam2d ? lam2d[(long long)ky * nx + kx] : lamx[kx] + lamy[ky];
      keep[i] = conjg(scale(czt_out(lds, py, ky), spec_mult<T, 0>(lam, alpha, beta, nrep, s)));
    }
  }
  czt_reload(lds, py, keep);
  czt_conv(lds, py, tw);
  for (int ky = threadIdx.x; ky < ny; ky += blockDim.x) col[(long long)ky * cols] = conjg(czt_out(lds, py, ky));
}

}  // namespace mjx
