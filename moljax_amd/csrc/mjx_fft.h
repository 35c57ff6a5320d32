// Device-side FFT building blocks for gfx950 (CDNA4, wave64).
//
// One "line" (a row or a column of a 2-D grid) of N complex points is held by
// TN = N/P threads, P points per thread.  Thread j holds, in registers, the
// points with natural index j + TN*m (m = 0..P-1) both on input to the first
// stage and on output of the last stage of a Stockham autosort FFT.  Middle
// stages exchange through LDS.  Because input and output index sets coincide,
// a forward FFT, a pointwise operation and an inverse FFT chain in registers
// with no data movement in between.
//
// Radix-R stage (R | P), current sub-transform length Ns:
//   butterfly b reads x[b + r*N/R] (r < R), multiplies input r by
//   w^(r*(b mod Ns)) with w = exp(SIGN*2*pi*i/(Ns*R)), runs a length-R DFT and
//   writes output r to (b/Ns)*Ns*R + (b mod Ns) + r*Ns.
// Twiddles come from a table of exp(-2*pi*i*t/MJX_TWN) evaluated once in FP64
// on the device (sincospi), rounded once to FP32 for the FP32 table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define MJX_TWN 4096

namespace mjx {

template <typename T>
struct __align__(2 * sizeof(T)) cpx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T r, T i) {
  cpx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return mk(a.re * s, a.im * s); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return mk(a.re, -a.im); }
// SIGN*i*a : forward transforms use SIGN = -1, inverse SIGN = +1.
template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
  return SIGN < 0 ? mk(a.im, -a.re) : mk(-a.im, a.re);
}

// ---------------------------------------------------------------- small DFTs
template <int SIGN, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a0, cpx<T>& a1) {
  cpx<T> t = a0 - a1;
  a0 = a0 + a1;
  a1 = t;
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
  cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_si<SIGN>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

// Radix-8 as two radix-4 DFTs (even/odd inputs) and one radix-2 layer.
template <int SIGN, typename T>
__device__ __forceinline__ void dft8(cpx<T>* a) {
  dft4<SIGN>(a[0], a[2], a[4], a[6]);
  dft4<SIGN>(a[1], a[3], a[5], a[7]);
  const T c = (T)0.70710678118654752440;
  cpx<T> o0 = a[1];
  cpx<T> o1 = cmul(a[3], mk<T>(c, SIGN * c));
  cpx<T> o2 = mul_si<SIGN>(a[5]);
  cpx<T> o3 = cmul(a[7], mk<T>(-c, SIGN * c));
  cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + o1;
  a[5] = e1 - o1;
  a[2] = e2 + o2;
  a[6] = e2 - o2;
  a[3] = e3 + o3;
  a[7] = e3 - o3;
}

// Radix-16 as 4 x 4: n = 4*n1 + n2, k = k1 + 4*k2.
template <int SIGN, typename T>
__device__ __forceinline__ void dft16(cpx<T>* a) {
  dft4<SIGN>(a[0], a[4], a[8], a[12]);
  dft4<SIGN>(a[1], a[5], a[9], a[13]);
  dft4<SIGN>(a[2], a[6], a[10], a[14]);
  dft4<SIGN>(a[3], a[7], a[11], a[15]);
  const T C1 = (T)0.92387953251128675613;  // cos(pi/8)
  const T S1 = (T)0.38268343236508977173;  // sin(pi/8)
  const T c = (T)0.70710678118654752440;
  // a[n2 + 4*k1] *= w16^(n2*k1)
  a[5] = cmul(a[5], mk<T>(C1, SIGN * S1));     // 1
  a[9] = cmul(a[9], mk<T>(c, SIGN * c));       // 2
  a[13] = cmul(a[13], mk<T>(S1, SIGN * C1));   // 3
  a[6] = cmul(a[6], mk<T>(c, SIGN * c));       // 2
  a[10] = mul_si<SIGN>(a[10]);                 // 4
  a[14] = cmul(a[14], mk<T>(-c, SIGN * c));    // 6
  a[7] = cmul(a[7], mk<T>(S1, SIGN * C1));     // 3
  a[11] = cmul(a[11], mk<T>(-c, SIGN * c));    // 6
  a[15] = cmul(a[15], mk<T>(-C1, -SIGN * S1)); // 9
  dft4<SIGN>(a[0], a[1], a[2], a[3]);
  dft4<SIGN>(a[4], a[5], a[6], a[7]);
  dft4<SIGN>(a[8], a[9], a[10], a[11]);
  dft4<SIGN>(a[12], a[13], a[14], a[15]);
  // a[4*k1 + k2] holds X[k1 + 4*k2]: transpose the 4x4 register block.
  cpx<T> t;
  t = a[1]; a[1] = a[4]; a[4] = t;
  t = a[2]; a[2] = a[8]; a[8] = t;
  t = a[3]; a[3] = a[12]; a[12] = t;
  t = a[6]; a[6] = a[9]; a[9] = t;
  t = a[7]; a[7] = a[13]; a[13] = t;
  t = a[11]; a[11] = a[14]; a[14] = t;
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* a) {
  if constexpr (R == 2) {
    dft2<SIGN>(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4<SIGN>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    dft8<SIGN>(a);
  } else {
    static_assert(R == 16, "radix must be 2, 4, 8 or 16");
    dft16<SIGN>(a);
  }
}

// ------------------------------------------------------------------- plans
// Points per thread and radix sequence for each supported power-of-two length.
__host__ __device__ constexpr int plan_p(int n) {
  return n == 16 ? 4 : n == 32 ? 8 : n == 64 ? 8 : n == 512 ? 8 : 16;
}
__host__ __device__ constexpr int plan_nstages(int n) {
  return n <= 256 ? 2 : 3;
}
__host__ __device__ constexpr int plan_radix(int n, int s) {
  // 16: 4x4  32: 8x4  64: 8x8  128: 16x8  256: 16x16
  // 512: 8x8x8  1024: 16x16x4  2048: 16x16x8  4096: 16x16x16
  return n == 16 ? 4
       : n == 32 ? (s == 0 ? 8 : 4)
       : n == 64 ? 8
       : n == 128 ? (s == 0 ? 16 : 8)
       : n == 256 ? 16
       : n == 512 ? 8
       : n == 1024 ? (s < 2 ? 16 : 4)
       : n == 2048 ? (s < 2 ? 16 : 8)
       : 16;
}
__host__ __device__ constexpr int plan_ns_before(int n, int s) {
  return s == 0 ? 1 : plan_ns_before(n, s - 1) * plan_radix(n, s - 1);
}
__host__ __device__ constexpr bool supported_len(int n) {
  return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 ||
         n == 4096;
}

template <int I, int E>
struct SFor {
  template <class F>
  __device__ __forceinline__ static void run(F&& f) {
    f(std::integral_constant<int, I>{});
    SFor<I + 1, E>::run(f);
  }
};
template <int E>
struct SFor<E, E> {
  template <class F>
  __device__ __forceinline__ static void run(F&&) {}
};

// ----------------------------------------------------------------- twiddles
// tw: device table of MJX_TWN entries exp(-2*pi*i*t/MJX_TWN) in precision T,
// owned by the library (one per device), passed to every kernel.

// ---------------------------------------------------------- twiddle cache
// The twiddles a thread needs depend only on its lane index j, so a forward
// and an inverse FFT in the same thread share one set of table loads:
// stage s >= 1, butterfly q, input r >= 1 uses w^(r * k), k = b mod Ns,
// b = j + TN*q (forward convention, conjugated for the inverse).
__host__ __device__ constexpr int tw_count_stage(int n, int s) {
  return (plan_p(n) / plan_radix(n, s)) * (plan_radix(n, s) - 1);
}
__host__ __device__ constexpr int tw_offset(int n, int s) {
  return s <= 1 ? 0 : tw_offset(n, s - 1) + tw_count_stage(n, s - 1);
}
__host__ __device__ constexpr int tw_count(int n) { return tw_offset(n, plan_nstages(n)); }

template <typename T, int N>
struct TwCache {
  cpx<T> w[tw_count(N)];
};

template <typename T, int N>
__device__ __forceinline__ void load_twiddles(TwCache<T, N>& c, int j, const cpx<T>* __restrict__ tw) {
  constexpr int P = plan_p(N);
  constexpr int TN = N / P;
  constexpr int NS = plan_nstages(N);
  SFor<1, NS>::run([&](auto S_) {
    constexpr int s = decltype(S_)::value;
    constexpr int R = plan_radix(N, s);
    constexpr int Ns = plan_ns_before(N, s);
    constexpr int NB = P / R;
    constexpr int TSTEP = MJX_TWN / (Ns * R);
    constexpr int off = tw_offset(N, s);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int k = (j + TN * q) & (Ns - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) c.w[off + q * (R - 1) + (r - 1)] = tw[r * k * TSTEP];
    }
  });
}

// LDS address of natural index idx: idx * ES, or, with PAD, one spare slot
// after every 16 elements (idx + idx/16) so that the stride-R writes of a
// radix-R stage spread over the LDS banks.
template <int ES, int PAD>
__device__ __forceinline__ int lds_addr(int idx) {
  return (PAD ? idx + (idx >> 4) : idx) * ES;
}
__host__ __device__ constexpr int lds_line_len(int n, int pad) { return pad ? n + n / 16 : n; }

// ------------------------------------------------------------- line FFT
// v     : P registers, element m <-> natural index j + TN*m (in and out)
// lds   : this line's LDS base; natural index idx lives at lds[lds_addr(idx)]
// twc   : this thread's twiddles (load_twiddles)
// The caller guarantees every thread of the block calls this (barriers).
template <typename T, int N, int SIGN, int ES, int PAD>
__device__ __forceinline__ void line_fft(cpx<T> (&v)[plan_p(N)], int j, cpx<T>* lds, const TwCache<T, N>& twc) {
  constexpr int P = plan_p(N);
  constexpr int TN = N / P;
  constexpr int NS = plan_nstages(N);
  SFor<0, NS>::run([&](auto S_) {
    constexpr int s = decltype(S_)::value;
    constexpr int R = plan_radix(N, s);
    constexpr int Ns = plan_ns_before(N, s);
    constexpr int NB = P / R;  // butterflies per thread
    static_assert(P % R == 0, "radix must divide points per thread");
    if constexpr (s > 0) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int b = j + TN * q;
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + r * NB] = lds[lds_addr<ES, PAD>(b + r * (N / R))];
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      cpx<T> t[R];
#pragma unroll
      for (int r = 0; r < R; ++r) t[r] = v[q + r * NB];
      if constexpr (s > 0) {
        constexpr int off = tw_offset(N, s);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cpx<T> w = twc.w[off + q * (R - 1) + (r - 1)];
          t[r] = cmul(t[r], SIGN < 0 ? w : conjg(w));
        }
      }
      dft<R, SIGN>(t);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * NB] = t[r];
    }
    if constexpr (s < NS - 1) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int b = j + TN * q;
        const int k = b & (Ns - 1);
        const int idxD = (b / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_addr<ES, PAD>(idxD + r * Ns)] = v[q + r * NB];
      }
      __syncthreads();
    }
  });
}

// Widest ETDRK4 column tile (columns per workgroup) for a precision and
// size: at most 1024 threads and 64 KiB of LDS per workgroup.
inline int etd_col_max(int dtype, int n) {
  const int tn = n / plan_p(n);
  const int cm0 = 1024 / tn < n / 2 ? 1024 / tn : n / 2;
  const int cl = 65536 / (n * (dtype == 1 ? 16 : 8));
  return cm0 < cl ? cm0 : cl;
}

}  // namespace mjx
