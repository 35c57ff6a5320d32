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
// on the device (sincospi), rounded once to FP32 for the FP32 table.  The
// mixed-radix lengths (stages of radix 5, 6, 10 and 12 beside the
// power-of-two ones) index a table of their own, exp(-2*pi*i*t/N) for t < N,
// stored after the power-of-two table (tw_base).
// A radix-R stage needs R | P (every thread runs P/R whole butterflies), so
// a length's points per thread P is a multiple of each of its radices.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define MJX_TWN 4096

// Lengths that are not powers of two but have a Stockham plan of their own
// (radices 2, 4, 5, 6, 10, 12); every other length runs the chirp-z kernels.
#define MJX_FOR_MIXED(X) X(96) X(100) X(192) X(200) X(250) X(384) X(500) X(768) X(1000)

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

// Radix-3: w = exp(SIGN 2 pi i / 3) = -1/2 + SIGN i sqrt(3)/2.
template <int SIGN, typename T>
__device__ __forceinline__ void dft3(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2) {
  const T h = (T)0.86602540378443864676;  // sqrt(3)/2
  const cpx<T> t = a1 + a2;
  const cpx<T> m = a0 - scale(t, (T)0.5);
  const cpx<T> d = mul_si<SIGN>(scale(a1 - a2, h));
  a0 = a0 + t;
  a1 = m + d;
  a2 = m - d;
}

// Radix-5 from the symmetric and antisymmetric input pairs:
// X1,4 = a1 +- SIGN i b1, X2,3 = a2 +- SIGN i b2 with
// a1 = x0 + c1 (x1+x4) + c2 (x2+x3), b1 = s1 (x1-x4) + s2 (x2-x3),
// a2 = x0 + c2 (x1+x4) + c1 (x2+x3), b2 = s2 (x1-x4) - s1 (x2-x3),
// c_k = cos(2 pi k / 5), s_k = sin(2 pi k / 5).
template <int SIGN, typename T>
__device__ __forceinline__ void dft5(cpx<T>* a) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  const cpx<T> t1 = a[1] + a[4], t2 = a[2] + a[3], t3 = a[1] - a[4], t4 = a[2] - a[3];
  const cpx<T> a1 = a[0] + scale(t1, c1) + scale(t2, c2);
  const cpx<T> a2 = a[0] + scale(t1, c2) + scale(t2, c1);
  const cpx<T> b1 = mul_si<SIGN>(scale(t3, s1) + scale(t4, s2));
  const cpx<T> b2 = mul_si<SIGN>(scale(t3, s2) - scale(t4, s1));
  a[0] = a[0] + t1 + t2;
  a[1] = a1 + b1;
  a[4] = a1 - b1;
  a[2] = a2 + b2;
  a[3] = a2 - b2;
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* a);

// Prime-factor (Good-Thomas) DFT of length N1*N2 with gcd(N1, N2) = 1: input
// n = (N2 n1 + N1 n2) mod N, N1-point DFTs over n1, N2-point DFTs over n2, output k = (N2 e1 k1 + N1 e2 k2) mod N with e1 = N2^-1 mod N1 and
// e2 = N1^-1 mod N2; W_N^(nk) = W_N1^(n1 k1) W_N2^(n2 k2), so no twiddles.
__host__ __device__ constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}
template <int N1, int N2, int SIGN, typename T>
__device__ __forceinline__ void dft_pfa(cpx<T>* a) {
  constexpr int N = N1 * N2;
  constexpr int E1 = inv_mod(N2 % N1, N1), E2 = inv_mod(N1 % N2, N2);
  cpx<T> y[N1][N2];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    cpx<T> t[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) t[n1] = a[(N2 * n1 + N1 * n2) % N];
    dft<N1, SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) y[k1][n2] = t[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    dft<N2, SIGN>(y[k1]);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) a[(N2 * E1 * k1 + N1 * E2 * k2) % N] = y[k1][k2];
  }
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* a) {
  if constexpr (R == 2) {
    dft2<SIGN>(a[0], a[1]);
  } else if constexpr (R == 3) {
    dft3<SIGN>(a[0], a[1], a[2]);
  } else if constexpr (R == 5) {
    dft5<SIGN>(a);
  } else if constexpr (R == 6) {
    dft_pfa<2, 3, SIGN>(a);
  } else if constexpr (R == 10) {
    dft_pfa<2, 5, SIGN>(a);
  } else if constexpr (R == 12) {
    dft_pfa<4, 3, SIGN>(a);
  } else if constexpr (R == 4) {
    dft4<SIGN>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    dft8<SIGN>(a);
  } else {
    static_assert(R == 16, "radix must be 2, 3, 4, 5, 6, 8, 10, 12 or 16");
    dft16<SIGN>(a);
  }
}

// ------------------------------------------------------------------- plans
__host__ __device__ constexpr bool pow2_len(int n) {
  return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 ||
         n == 4096;
}
__host__ __device__ constexpr bool mixed_len(int n) {
  return n == 96 || n == 100 || n == 192 || n == 200 || n == 250 || n == 384 || n == 500 || n == 768 ||
         n == 1000;
}
__host__ __device__ constexpr bool supported_len(int n) { return pow2_len(n) || mixed_len(n); }

// Points per thread and radix sequence for each supported length.
//   power of two: 16: 4x4  32: 8x4  64: 8x8  128: 16x8  256: 16x16
//                 512: 8x8x8  1024: 16x16x4  2048: 16x16x8  4096: 16x16x16
//   mixed (P):    96: 4x4x6 (12)  100: 10x10 (10)  192: 4x4x12 (12)
//                 200: 10x10x2 (10)  250: 10x5x5 (10)  384: 4x4x12x2 (12)
//                 500: 10x10x5 (10)  768: 4x4x4x12 (12)  1000: 10x10x10 (10)
// (radices 6, 10 and 12 are prime-factor 2x3, 2x5 and 4x3 DFTs).
__host__ __device__ constexpr int plan_p(int n) {
  return n == 16 ? 4 : n == 32 ? 8 : n == 64 ? 8 : n == 512 ? 8
       : (n == 96 || n == 192 || n == 384 || n == 768) ? 12
       : mixed_len(n) ? 10
       : 16;
}
__host__ __device__ constexpr int plan_nstages(int n) {
  return n == 100 ? 2
       : (n == 96 || n == 192 || n == 200 || n == 250 || n == 500 || n == 1000) ? 3
       : (n == 384 || n == 768) ? 4
       : n <= 256 ? 2 : 3;
}
__host__ __device__ constexpr int plan_radix(int n, int s) {
  return n == 16 ? 4
       : n == 32 ? (s == 0 ? 8 : 4)
       : n == 64 ? 8
       : n == 128 ? (s == 0 ? 16 : 8)
       : n == 256 ? 16
       : n == 512 ? 8
       : n == 1024 ? (s < 2 ? 16 : 4)
       : n == 2048 ? (s < 2 ? 16 : 8)
       : n == 4096 ? 16
       : n == 96 ? (s < 2 ? 4 : 6)
       : (n == 100 || n == 1000) ? 10
       : n == 192 ? (s < 2 ? 4 : 12)
       : n == 200 ? (s < 2 ? 10 : 2)
       : n == 250 ? (s == 0 ? 10 : 5)
       : n == 384 ? (s < 2 ? 4 : s == 2 ? 12 : 2)
       : n == 500 ? (s < 2 ? 10 : 5)
       : (s < 3 ? 4 : 12);  // 768
}
__host__ __device__ constexpr int plan_ns_before(int n, int s) {
  return s == 0 ? 1 : plan_ns_before(n, s - 1) * plan_radix(n, s - 1);
}

// Twiddle table of a length: the shared power-of-two table (MJX_TWN entries)
// or the length's own N entries, stored after it in the order of
// MJX_FOR_MIXED.
__host__ __device__ constexpr int tw_len(int n) { return mixed_len(n) ? n : MJX_TWN; }
__host__ __device__ constexpr int tw_base(int n) {
  return !mixed_len(n) ? 0
       : MJX_TWN + (n > 96 ? 96 : 0) + (n > 100 ? 100 : 0) + (n > 192 ? 192 : 0) + (n > 200 ? 200 : 0) +
             (n > 250 ? 250 : 0) + (n > 384 ? 384 : 0) + (n > 500 ? 500 : 0) + (n > 768 ? 768 : 0);
}
constexpr int MJX_TW_TOTAL = MJX_TWN + 96 + 100 + 192 + 200 + 250 + 384 + 500 + 768 + 1000;

// Compile-time checks of every plan: the radices multiply to N and divide P.
__host__ __device__ constexpr bool plan_ok(int n) {
  int prod = 1;
  for (int s = 0; s < plan_nstages(n); ++s) {
    if (plan_p(n) % plan_radix(n, s) != 0) return false;
    prod *= plan_radix(n, s);
  }
  return prod == n && n % plan_p(n) == 0;
}
#define MJX_PLAN_CHECK(N) static_assert(plan_ok(N), "bad FFT plan for " #N);
MJX_FOR_MIXED(MJX_PLAN_CHECK)
MJX_PLAN_CHECK(16) MJX_PLAN_CHECK(32) MJX_PLAN_CHECK(64) MJX_PLAN_CHECK(128) MJX_PLAN_CHECK(256)
MJX_PLAN_CHECK(512) MJX_PLAN_CHECK(1024) MJX_PLAN_CHECK(2048) MJX_PLAN_CHECK(4096)
#undef MJX_PLAN_CHECK
static_assert(tw_base(1000) + 1000 == MJX_TW_TOTAL, "mixed twiddle tables overlap");

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

// The mixed-radix plans whose twiddles would take more than 128 bytes of
// registers per thread (32 FP32 or 16 FP64 values) read them from the table
// (L1-resident) at each use instead.
template <typename T, int N, bool USE = true>
struct TwCache {
  // USE = false: always read the table at each use (kernels that need the
  // registers for other operands)
  static constexpr bool CACHED = USE && (pow2_len(N) || tw_count(N) * (int)sizeof(T) <= 128);
  cpx<T> w[CACHED ? tw_count(N) : 1];
  const cpx<T>* tw;
};

template <typename T, int N, bool USE>
__device__ __forceinline__ void load_twiddles(TwCache<T, N, USE>& c, int j, const cpx<T>* __restrict__ tw) {
  c.tw = tw;
  if constexpr (!TwCache<T, N, USE>::CACHED) return;
  constexpr int P = plan_p(N);
  constexpr int TN = N / P;
  constexpr int NS = plan_nstages(N);
  SFor<1, NS>::run([&](auto S_) {
    constexpr int s = decltype(S_)::value;
    constexpr int R = plan_radix(N, s);
    constexpr int Ns = plan_ns_before(N, s);
    constexpr int NB = P / R;
    constexpr int TSTEP = tw_len(N) / (Ns * R);
    static_assert(TSTEP * Ns * R == tw_len(N), "twiddle table step");
    constexpr int off = tw_offset(N, s);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int k = pow2_len(N) ? ((j + TN * q) & (Ns - 1)) : (int)((unsigned)(j + TN * q) % (unsigned)Ns);
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
template <typename T, int N, int SIGN, int ES, int PAD, bool USE>
__device__ __forceinline__ void line_fft(cpx<T> (&v)[plan_p(N)], int j, cpx<T>* lds, const TwCache<T, N, USE>& twc) {
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
          cpx<T> w;
          if constexpr (TwCache<T, N, USE>::CACHED) {
            w = twc.w[off + q * (R - 1) + (r - 1)];
          } else {
            constexpr int TSTEP = tw_len(N) / (Ns * R);
            const int k = (int)((unsigned)(j + TN * q) % (unsigned)Ns);
            w = twc.tw[r * k * TSTEP];
          }
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
        int k, idxD;
        if constexpr (pow2_len(N)) {
          k = b & (Ns - 1);
          idxD = (b / Ns) * Ns * R + k;
        } else {
          k = (int)((unsigned)b % (unsigned)Ns);
          idxD = (int)((unsigned)b / (unsigned)Ns) * Ns * R + k;
        }
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
