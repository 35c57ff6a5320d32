// Real-to-real line transforms (DST-I for homogeneous Dirichlet, DCT-II /
// DCT-III for homogeneous Neumann) built on the complex line FFT, and the
// spectral solves that use them.
//
// Two real lines a, b share one complex FFT of length L:
//   DST-I (N = L/2 - 1 points, M = N + 1): z = oddext(a) + i*oddext(b),
//     oddext(x) = [0, x_0..x_{N-1}, 0, -x_{N-1}..-x_0];  Z = FFT(z) gives
//     DST(a)_k = -Im(Z_k)/2, DST(b)_k = Re(Z_k)/2 for k = 1..N, with
//     DST(x)_k = sum_n x_n sin(pi k (n+1)/M) (moljax/core/fft_nonperiodic.py:97-122).
//     Z is odd in k, so a per-k scaling of the two real spectra is a
//     componentwise scaling of (Re Z, Im Z) and one inverse FFT gives both
//     (2/M) DST(m DST(x)) as Re/(2M), Im/(2M).
//   DCT-II (N = L/2 points): z = evenext(a) + i*evenext(b),
//     evenext(x) = [x_0..x_{N-1}, x_{N-1}..x_0];  with th_k = pi k/(2N),
//     Q_k = e^{-i th_k} Z_k / 2 = C(a)_k + i C(b)_k for k < N (and
//     Q_k = e^{+i th_k} Z_{2N-k} / 2), C(x)_k = sum_n x_n cos(pi k (2n+1)/(2N)).
//     The unnormalised DCT-III x_n = sum_k g_k cos(pi k (2n+1)/(2N)) of two
//     real sequences is the inverse FFT of the Hermitian Z'_e = e^{i th_e} Q'_e
//     (e < N), Z'_0 = 2 Q'_0, Z'_N = 0, Z'_{2N-k} = e^{-i th_k} Q'_k, halved.
// The ortho DCT of the reference (jax.scipy.fft.dct norm='ortho',
// fft_nonperiodic.py:352-355) is the unnormalised pair with f_k^2 = 1/N (k=0),
// 2/N (k>0) folded into the spectral multiplier.
//   DST-II (cell-centred Dirichlet, ghost = -interior as moljax/core/bc.py
//     :237-258 fills it): S(x)_k = sum_n x_n sin(pi (k+1)(2n+1)/(2N)) equals
//     C((-1)^n x_n)_{N-1-k}, since sin(pi(n+1/2) - t) = (-1)^n cos t.  With
//     `alt` set, the DCT path alternates the sign along the line on input
//     (forward, apply) and output (inverse, apply) and reads the line's
//     symbol in reverse order, so the DCT-II machinery diagonalises the
//     Dirichlet stencil; the forward transform stores its output reversed
//     (DST-II order) and the inverse (DST-III) reads its input reversed.
// A boundary lift (bc.py ghost values folded into the right-hand side) is
// added while the first pass loads its lines: lift = {left, right, bottom,
// top} contributions of the ghost ring to the 5-point Laplacian, scaled by
// beta (Helmholtz) or -1 (Poisson) per instance.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "mjx_common.h"
#include "mjx_fft.h"

namespace mjx {

enum { R2R_DST = 0, R2R_DCT = 1 };
enum { R2R_FWD = 0, R2R_INV = 1, R2R_APPLY = 2 };
enum { R2R_ROW = 0, R2R_COL = 1 };
enum { R2R_HELM = 0, R2R_POISSON = 1, R2R_HELM_POW = 2, R2R_EXP = 3, R2R_PHI1 = 4 };

const cpx<float>* twiddles_f32();
const cpx<double>* twiddles_f64();

struct R2RArgs {
  const void* in;
  long long in_es, in_ls, in_is;  // element, line, instance strides (elements)
  void* out;
  long long out_es, out_ls, out_is;
  int nlines;                     // real lines per instance
  int tiles;                      // line tiles per instance
  const void* lam_line;           // symbol along the transform (N entries)
  const void* lam_cross;          // symbol of the other dimension per line (or null)
  int n_cross;                    // other dimension's length (DCT normalisation), 0 = none
  const double* inst;             // per-instance rows (beta_a, -, alpha_a, ...)
  int mmode;                      // R2R_HELM or R2R_POISSON
  double scale;                   // extra output scale (FWD / INV)
  int alt;                        // DCT path as DST-II / DST-III (cell-centred Dirichlet)
  int accumulate;                 // add the result to the output instead of storing it
  double beta0;                   // beta when inst is null (alpha 0, power 1)
  int lift_on;                    // add the boundary lift while loading
  double lift[4];                 // left, right (line ends), bottom, top (first / last line)
};

template <int KIND, int L>
struct R2RCfg {
  static constexpr int N = KIND == R2R_DST ? L / 2 - 1 : L / 2;
  static constexpr int M = L / 2;
  static constexpr int P = plan_p(L);
  static constexpr int TN = L / P;
  static constexpr int LP = lds_line_len(L, 1);
  static constexpr int TW = N + 1;  // padded tile row
};

// e^{i*sgn*pi*k/(2N)} for the DCT twiddle, from the table exp(-2 pi i t/MJX_TWN).
template <typename T, int N, int SGN>
__device__ __forceinline__ cpx<T> dct_rot(const cpx<T>* __restrict__ tw, int k) {
  const cpx<T> w = tw[k * (MJX_TWN / (4 * N))];
  return SGN > 0 ? conjg(w) : w;
}

// Helmholtz ((1 + alpha lam) / (1 - beta lam)), raised to `power` in
// R2R_HELM_POW mode (n backward-Euler / Crank-Nicolson steps in one solve),
// or Poisson 1 / lam.
// R2R_EXP: e^{beta lam}; R2R_PHI1: beta phi1(beta lam) with the quadratic
// series below |z| = 1e-4 (moljax/core/fft_nonperiodic.py:381-389).
template <typename T>
__device__ __forceinline__ T r2r_mult(T lam, T alpha, T beta, int mmode, int power) {
  if (mmode == R2R_EXP) return exp(beta * lam);
  if (mmode == R2R_PHI1) {
    const T z = beta * lam;
    return beta * (fabs(z) < (T)1e-4 ? (T)1 + z / (T)2 + z * z / (T)6 : (exp(z) - (T)1) / z);
  }
  if (mmode == R2R_POISSON) {
    if (fabs(lam) < (T)1e-14) return (T)0;
    if constexpr (std::is_same<T, double>::value) return 1.0 / lam;
    else return 1.0f / lam;
  }
  const T g = ((T)1 + alpha * lam) / ((T)1 - beta * lam);
  if (mmode != R2R_HELM_POW) return g;
  T m = (T)1, b = g;
  for (int n = power; n > 0; n >>= 1) {
    if (n & 1) m *= b;
    b *= b;
  }
  return m;
}

template <typename T, int KIND, int L, int ORIENT, int MODE, int TL>
__global__ __launch_bounds__((TL * R2RCfg<KIND, L>::TN)) void k_r2r(R2RArgs a, const cpx<T>* __restrict__ tw) {
  using C = R2RCfg<KIND, L>;
  constexpr int N = C::N, M = C::M, P = C::P, TN = C::TN, TW = C::TW;
  constexpr int NL = 2 * TL;  // real lines per block
  __shared__ cpx<T> lfft[TL * C::LP];
  __shared__ T tile[NL * TW];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / a.tiles;
  const int line0 = (w - inst * a.tiles) * NL;
  const int tid = threadIdx.x;
  const T* in = (const T*)a.in + inst * a.in_is;
  T* out = (T*)a.out + inst * a.out_is;

  for (int idx = tid; idx < NL * N; idx += TL * TN) {
    int l, e;
    if constexpr (ORIENT == R2R_ROW) {
      l = idx / N;
      e = idx - l * N;
    } else {
      l = idx % NL;
      e = idx / NL;
    }
    const int g = line0 + l;
    T x = g < a.nlines ? in[(long long)g * a.in_ls + (long long)e * a.in_es] : (T)0;
    int es = e;
    if constexpr (KIND == R2R_DCT) {
      if (a.lift_on && g < a.nlines) {
        const double w = a.mmode == R2R_POISSON ? -1.0 : a.inst[8 * inst];
        const double add = (e == 0 ? a.lift[0] : 0.0) + (e == N - 1 ? a.lift[1] : 0.0) +
                           (g == 0 ? a.lift[2] : 0.0) + (g == a.nlines - 1 ? a.lift[3] : 0.0);
        x += (T)(w * add);
      }
      if (a.alt) {
        if constexpr (MODE == R2R_INV) es = N - 1 - e;
        else if (e & 1) x = -x;
      }
    }
    tile[l * TW + es] = x;
  }
  __syncthreads();

  const int t = tid / TN, j = tid % TN;
  const T* ta = tile + (2 * t) * TW;
  const T* tb = tile + (2 * t + 1) * TW;
  cpx<T>* lline = lfft + t * C::LP;
  TwCache<T, L> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[P];

  if constexpr (KIND == R2R_DCT && MODE == R2R_INV) {
    // Hermitian spectrum of two unnormalised DCT-III inputs, then inverse FFT.
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int e = j + TN * m;
      cpx<T> z = mk((T)0, (T)0);
      if (e < N) {
        const cpx<T> q = mk(ta[e], tb[e]);
        z = e == 0 ? scale(q, (T)2) : cmul(dct_rot<T, N, +1>(tw, e), q);
      } else if (e > N) {
        const int k = L - e;
        z = cmul(dct_rot<T, N, -1>(tw, k), mk(ta[k], tb[k]));
      }
      v[m] = z;
    }
    line_fft<T, L, +1, 1, 1>(v, j, lline, twc);
  } else {
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int e = j + TN * m;
      T xa = (T)0, xb = (T)0;
      if constexpr (KIND == R2R_DST) {
        if (e >= 1 && e <= N) {
          xa = ta[e - 1];
          xb = tb[e - 1];
        } else if (e > M) {
          xa = -ta[L - 1 - e];
          xb = -tb[L - 1 - e];
        }
      } else {
        const int s = e < N ? e : L - 1 - e;
        xa = ta[s];
        xb = tb[s];
      }
      v[m] = mk(xa, xb);
    }
    line_fft<T, L, -1, 1, 1>(v, j, lline, twc);

    if constexpr (MODE == R2R_APPLY) {
      const double* ip = a.inst ? a.inst + 8 * inst : nullptr;
      const T beta = ip ? (T)ip[0] : (T)a.beta0, alpha = ip ? (T)ip[2] : (T)0;
      const int power = (ip && a.mmode == R2R_HELM_POW) ? (int)ip[7] : 1;
      const T* lam_line = (const T*)a.lam_line;
      const T* lam_cross = (const T*)a.lam_cross;
      const int ga = line0 + 2 * t, gb = ga + 1;
      // The cross dimension arrives in stored order: DCT index order, or
      // DST-II order (reversed DCT index) with alt; the symbols follow the
      // stored order and the DCT normalisation the DCT index.
      const T ca = lam_cross ? lam_cross[ga < a.nlines ? ga : 0] : (T)0;
      const T cb = lam_cross ? lam_cross[gb < a.nlines ? gb : 0] : (T)0;
      T fa = (T)1, fb = (T)1;
      if constexpr (KIND == R2R_DCT) {
        if (a.n_cross > 0) {
          const int da = a.alt ? a.n_cross - 1 - ga : ga, db = a.alt ? a.n_cross - 1 - gb : gb;
          fa = (T)((da == 0 ? 1.0 : 2.0) / a.n_cross);
          fb = (T)((db == 0 ? 1.0 : 2.0) / a.n_cross);
        }
      }
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const int e = j + TN * m;
        if constexpr (KIND == R2R_DST) {
          const int k = e <= M ? e : L - e;
          if (k >= 1 && k <= N) {
            const T ll = lam_line[k - 1];
            const T ma = r2r_mult<T>(ll + ca, alpha, beta, a.mmode, power);
            const T mb = r2r_mult<T>(ll + cb, alpha, beta, a.mmode, power);
            v[m] = mk(mb * v[m].re, ma * v[m].im);
          } else {
            v[m] = mk((T)0, (T)0);
          }
        } else {
          if (e == N) {
            v[m] = mk((T)0, (T)0);
          } else {
            const int k = e < N ? e : L - e;
            const cpx<T> q = e < N ? cmul(dct_rot<T, N, -1>(tw, k), v[m]) : cmul(dct_rot<T, N, +1>(tw, k), v[m]);
            const T f2 = (T)((k == 0 ? 1.0 : 2.0) / N);
            const T ll = lam_line[a.alt ? N - 1 - k : k];
            const T ma = r2r_mult<T>(ll + ca, alpha, beta, a.mmode, power) * f2 * fa;
            const T mb = r2r_mult<T>(ll + cb, alpha, beta, a.mmode, power) * f2 * fb;
            // Q'/1 with the 1/2 of Q = e^{-i th} Z / 2 and the doubled DC
            // term folded in: Z' = e^{+i th} Q' (e < N), 2 Q'_0 at e = 0.
            const cpx<T> qp = mk((T)0.5 * ma * q.re, (T)0.5 * mb * q.im);
            v[m] = e == 0 ? scale(qp, (T)2) : (e < N ? cmul(dct_rot<T, N, +1>(tw, k), qp)
                                                     : cmul(dct_rot<T, N, -1>(tw, k), qp));
          }
        }
      }
      line_fft<T, L, +1, 1, 1>(v, j, lline, twc);
    }
  }

  // Extract outputs into the tile (line 2t <- a, 2t+1 <- b).
  T* oa = tile + (2 * t) * TW;
  T* ob = tile + (2 * t + 1) * TW;
  const T sc = (T)a.scale;
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int e = j + TN * m;
    if constexpr (KIND == R2R_DST) {
      if (e >= 1 && e <= N) {
        if constexpr (MODE == R2R_APPLY) {
          oa[e - 1] = v[m].re * (T)(0.5 / M);
          ob[e - 1] = v[m].im * (T)(0.5 / M);
        } else {
          oa[e - 1] = -v[m].im * (T)0.5 * sc;
          ob[e - 1] = v[m].re * (T)0.5 * sc;
        }
      }
    } else {
      if (e < N) {
        if constexpr (MODE == R2R_FWD) {
          const cpx<T> q = cmul(dct_rot<T, N, -1>(tw, e), v[m]);
          oa[e] = q.re * (T)0.5 * sc;
          ob[e] = q.im * (T)0.5 * sc;
        } else {
          oa[e] = v[m].re * (T)0.5 * sc;
          ob[e] = v[m].im * (T)0.5 * sc;
        }
      }
    }
  }
  __syncthreads();
  for (int idx = tid; idx < NL * N; idx += TL * TN) {
    int l, e;
    if constexpr (ORIENT == R2R_ROW) {
      l = idx / N;
      e = idx - l * N;
    } else {
      l = idx % NL;
      e = idx / NL;
    }
    const int g = line0 + l;
    if (g < a.nlines) {
      T y = tile[l * TW + e];
      if constexpr (KIND == R2R_DCT) {
        if (a.alt) {
          if constexpr (MODE == R2R_FWD) y = tile[l * TW + N - 1 - e];
          else if (e & 1) y = -y;
        }
      }
      T* o = out + (long long)g * a.out_ls + (long long)e * a.out_es;
      *o = a.accumulate ? *o + y : y;
    }
  }
}

template <typename T, int L>
constexpr int r2r_lines() {
  constexpr int bt = sizeof(T) == 8 ? 128 : 256;
  return (L / plan_p(L)) >= bt ? 1 : bt / (L / plan_p(L));
}

// Lines per workgroup: the default (about 256 threads in FP32, 128 in
// FP64), or a quarter of it when the default would leave most CUs idle (a
// single 512^2 instance gives 64 workgroups): each pair of real lines shares
// one complex transform either way, so the results are the same bitwise.
template <typename T, int KIND, int L, int ORIENT, int MODE>
int r2r_launch(R2RArgs a, int B, hipStream_t st) {
  constexpr int TL = r2r_lines<T, L>();
  constexpr int TLS = TL >= 4 ? TL / 4 : 1;
  const cpx<T>* tw;
  if constexpr (std::is_same<T, float>::value) tw = twiddles_f32();
  else tw = twiddles_f64();
  const long long wide = (long long)((a.nlines + 2 * TL - 1) / (2 * TL)) * B;
  if (TLS < TL && wide < 256) {
    a.tiles = (a.nlines + 2 * TLS - 1) / (2 * TLS);
    hipLaunchKernelGGL((k_r2r<T, KIND, L, ORIENT, MODE, TLS>), dim3(a.tiles * B), dim3(TLS * R2RCfg<KIND, L>::TN),
                       0, st, a, tw);
    return check_launch("r2r");
  }
  a.tiles = (a.nlines + 2 * TL - 1) / (2 * TL);
  hipLaunchKernelGGL((k_r2r<T, KIND, L, ORIENT, MODE, TL>), dim3(a.tiles * B), dim3(TL * R2RCfg<KIND, L>::TN), 0,
                     st, a, tw);
  return check_launch("r2r");
}

#define MJX_FOR_R2R_LENGTHS(X) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

// Runtime dispatch on the line length L = 2(N+1) (DST) or 2N (DCT).
template <typename T>
int r2r_dispatch(int kind, int orient, int mode, int L, R2RArgs a, int B, hipStream_t st) {
#define X(LL)                                                                                        \
  if (L == LL) {                                                                                     \
    if (kind == R2R_DST) {                                                                           \
      if (orient == R2R_ROW) {                                                                       \
        if (mode == R2R_FWD) return r2r_launch<T, R2R_DST, LL, R2R_ROW, R2R_FWD>(a, B, st);           \
        if (mode == R2R_INV) return r2r_launch<T, R2R_DST, LL, R2R_ROW, R2R_INV>(a, B, st);           \
        return r2r_launch<T, R2R_DST, LL, R2R_ROW, R2R_APPLY>(a, B, st);                              \
      }                                                                                              \
      return r2r_launch<T, R2R_DST, LL, R2R_COL, R2R_APPLY>(a, B, st);                                \
    }                                                                                                \
    if (orient == R2R_ROW) {                                                                         \
      if (mode == R2R_FWD) return r2r_launch<T, R2R_DCT, LL, R2R_ROW, R2R_FWD>(a, B, st);             \
      if (mode == R2R_INV) return r2r_launch<T, R2R_DCT, LL, R2R_ROW, R2R_INV>(a, B, st);             \
      return r2r_launch<T, R2R_DCT, LL, R2R_ROW, R2R_APPLY>(a, B, st);                                \
    }                                                                                                \
    return r2r_launch<T, R2R_DCT, LL, R2R_COL, R2R_APPLY>(a, B, st);                                  \
  }
  MJX_FOR_R2R_LENGTHS(X)
#undef X
  return fail(MJX_E_SIZE, "unsupported transform length " + std::to_string(L));
}

}  // namespace mjx
