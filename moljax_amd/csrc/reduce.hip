// Deterministic FP64-accumulated reductions over strided 2-D interiors
// (the state norms of moljax/core/state.py:115-189,312-343 and is_finite of
// moljax/core/utils.py:186-197).
//
// Two launches per call: k_reduce_partial gives every instance a fixed set
// of RED_BLOCKS workgroups; each walks its rows of the instance in a fixed
// order (16-byte loads where the rows allow), reduces its 256 lanes by
// wave64 shuffles and LDS in a fixed tree, and writes one partial.  k_reduce_final folds an instance's partials with
// one wave, again in a fixed order, and combines the result with out[b]
// when accumulating over fields.  The same inputs therefore give bitwise
// the same result on every run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdint>

#include "mjx_common.h"
#include "moljax_hip.h"

namespace mjx {
namespace {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 64;

__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }

template <int OP>
__device__ __forceinline__ double combine(double a, double b) {
  if constexpr (OP == MJX_RED_MAXABS) return nanmax(a, b);
  else return a + b;
}

template <int OP>
__device__ __forceinline__ double wave_reduce(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = combine<OP>(v, __shfl_xor(v, o, 64));
  return v;
}

template <typename T, int OP>
__device__ __forceinline__ double term(T a, T b, double atol, double rtol) {
  const double x = (double)a;
  if constexpr (OP == MJX_RED_SUMSQ) return x * x;
  else if constexpr (OP == MJX_RED_DOT) return x * (double)b;
  else if constexpr (OP == MJX_RED_MAXABS) return fabs(x);
  else if constexpr (OP == MJX_RED_SUM) return x;
  else if constexpr (OP == MJX_RED_SCALED_SQ) {
    const double q = x / (atol + rtol * fabs((double)b));
    return q * q;
  } else return isfinite(x) ? 0.0 : 1.0;  // MJX_RED_NONFINITE
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
};
template <>
struct Vec16<double> {
  using type = double2;
};

// One wave per row: rows blk*4 + wave, then every 4 * nblk rows (nblk
// workgroups per instance, at most RED_BLOCKS).  A
// row is read as an unaligned head of scalars, a body of 16-byte loads (one
// per lane per step) and a tail of scalars; when b's row is not aligned like
// a's, the row is read one element per lane.  The combination order depends
// only on the shapes and on the rows' 16-byte alignment.
template <typename T, int OP>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partial(const T* __restrict__ a, long long ap, long long ai,
                                                                const T* __restrict__ b, long long bp, long long bi,
                                                                int ny, int nx, double atol, double rtol, int nblk,
                                                                double* __restrict__ partial) {
  using V = typename Vec16<T>::type;
  constexpr int VE = 16 / sizeof(T);
  const int inst = blockIdx.x / nblk, blk = blockIdx.x - inst * nblk;
  __shared__ double lds[RED_THREADS / 64];
  const T* A = a + inst * ai;
  const T* Bp = b ? b + inst * bi : nullptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0.0;  // identity of + and of max over |x|
  const int ystep = nblk * (RED_THREADS / 64);
  int y = blk * (RED_THREADS / 64) + wave;
  // Rows whose body fits one 16-byte load per lane (nx <= 64 * VE, the
  // interior of a padded field) with b absent or aligned like a on every
  // row: RROWS rows at a time, all their loads issued before any is used.
  // Each row keeps its own unaligned head and tail (an odd pitch alternates
  // the rows' alignment).
  constexpr int RROWS = 4;
  const bool batched = nx <= 64 * VE &&
                       (!Bp || ((((uintptr_t)A ^ (uintptr_t)Bp) & 15) == 0 &&
                                ((uintptr_t)((ap - bp) * (long long)sizeof(T)) & 15) == 0));
  if (batched) {
    for (; y + (RROWS - 1) * ystep < ny; y += RROWS * ystep) {
      V av[RROWS], bv[RROWS];
      T ha[RROWS], hb[RROWS], ta[RROWS], tb[RROWS];
      int hh[RROWS], nbb[RROWS], tt[RROWS];
#pragma unroll
      for (int r = 0; r < RROWS; ++r) {
        const T* ra = A + (long long)(y + r * ystep) * ap;
        const T* rb = Bp ? Bp + (long long)(y + r * ystep) * bp : nullptr;
        int h = (int)(((16 - ((uintptr_t)ra & 15)) & 15) / sizeof(T));
        if (h > nx) h = nx;
        const int nb = (nx - h) / VE;
        hh[r] = h;
        nbb[r] = nb;
        tt[r] = h + nb * VE;
        if (lane < nb) av[r] = *reinterpret_cast<const V*>(ra + h + lane * VE);
        if (rb && lane < nb) bv[r] = *reinterpret_cast<const V*>(rb + h + lane * VE);
        if (lane < h) {
          ha[r] = ra[lane];
          hb[r] = rb ? rb[lane] : (T)0;
        }
        if (tt[r] + lane < nx) {
          ta[r] = ra[tt[r] + lane];
          tb[r] = rb ? rb[tt[r] + lane] : (T)0;
        }
      }
#pragma unroll
      for (int r = 0; r < RROWS; ++r) {
        if (lane < hh[r]) acc = combine<OP>(acc, term<T, OP>(ha[r], hb[r], atol, rtol));
        if (lane < nbb[r]) {
          const T* ae = reinterpret_cast<const T*>(&av[r]);
          const T* be = reinterpret_cast<const T*>(&bv[r]);
#pragma unroll
          for (int e = 0; e < VE; ++e) acc = combine<OP>(acc, term<T, OP>(ae[e], Bp ? be[e] : (T)0, atol, rtol));
        }
        if (tt[r] + lane < nx) acc = combine<OP>(acc, term<T, OP>(ta[r], tb[r], atol, rtol));
      }
    }
  }
  // remaining rows (and every row of other layouts) one at a time
  for (; y < ny; y += ystep) {
    const T* ra = A + (long long)y * ap;
    const T* rb = Bp ? Bp + (long long)y * bp : nullptr;
    const uintptr_t ua = (uintptr_t)ra;
    if (rb && ((ua ^ (uintptr_t)rb) & 15)) {
      for (int x = lane; x < nx; x += 64) acc = combine<OP>(acc, term<T, OP>(ra[x], rb[x], atol, rtol));
      continue;
    }
    int h = (int)(((16 - (ua & 15)) & 15) / sizeof(T));
    if (h > nx) h = nx;
    const int nb = (nx - h) / VE;
    const int t0 = h + nb * VE;
    if (lane < h) acc = combine<OP>(acc, term<T, OP>(ra[lane], rb ? rb[lane] : (T)0, atol, rtol));
    for (int v = lane; v < nb; v += 64) {
      const V av = *reinterpret_cast<const V*>(ra + h + v * VE);
      const T* ae = reinterpret_cast<const T*>(&av);
      if (rb) {
        const V bv = *reinterpret_cast<const V*>(rb + h + v * VE);
        const T* be = reinterpret_cast<const T*>(&bv);
#pragma unroll
        for (int e = 0; e < VE; ++e) acc = combine<OP>(acc, term<T, OP>(ae[e], be[e], atol, rtol));
      } else {
#pragma unroll
        for (int e = 0; e < VE; ++e) acc = combine<OP>(acc, term<T, OP>(ae[e], (T)0, atol, rtol));
      }
    }
    if (t0 + lane < nx) acc = combine<OP>(acc, term<T, OP>(ra[t0 + lane], rb ? rb[t0 + lane] : (T)0, atol, rtol));
  }
  acc = wave_reduce<OP>(acc);
  if (lane == 0) lds[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = lds[0];
    for (int w = 1; w < RED_THREADS / 64; ++w) s = combine<OP>(s, lds[w]);
    partial[(long long)inst * RED_BLOCKS + blk] = s;
  }
}

template <int OP>
__global__ __launch_bounds__(64) void k_reduce_final(const double* __restrict__ partial, double* __restrict__ out,
                                                     int accumulate, int nblk) {
  const int inst = blockIdx.x;
  // one partial per lane (RED_BLOCKS == 64); lanes past nblk hold the
  // identity (0 for every reduction here, including the max of |x|)
  double s = (int)threadIdx.x < nblk ? partial[(long long)inst * RED_BLOCKS + threadIdx.x] : 0.0;
  s = wave_reduce<OP>(s);
  if (threadIdx.x == 0) out[inst] = accumulate ? combine<OP>(out[inst], s) : s;
}

static_assert(RED_BLOCKS == 64, "k_reduce_final folds one partial per lane");

// Batch total of per-instance results in a fixed order (lane l folds
// instances l, l + 64, ..., then a wave tree), optionally its square root.
template <int OP>
__global__ __launch_bounds__(64) void k_reduce_total(const double* __restrict__ per_inst, int B, int take_sqrt,
                                                     double* __restrict__ total) {
  double s = 0.0;
  for (int b = threadIdx.x; b < B; b += 64) s = combine<OP>(s, per_inst[b]);
  s = wave_reduce<OP>(s);
  if (threadIdx.x == 0) total[0] = take_sqrt ? sqrt(s) : s;
}

template <typename T, int OP>
int launch(int B, int ny, int nx, const T* a, long long ap, long long ai, const T* b, long long bp, long long bi,
           double atol, double rtol, double* out, double* work, int accumulate, hipStream_t st) {
  // workgroups per instance: one per 32 rows (4 waves x 8 rows), at most
  // RED_BLOCKS
  const int nblk = std::min(RED_BLOCKS, std::max(1, (ny + 31) / 32));
  hipLaunchKernelGGL((k_reduce_partial<T, OP>), dim3(B * nblk), dim3(RED_THREADS), 0, st, a, ap, ai, b, bp, bi, ny,
                     nx, atol, rtol, nblk, work);
  int rc = check_launch("reduce_partial");
  if (rc) return rc;
  hipLaunchKernelGGL((k_reduce_final<OP>), dim3(B), dim3(64), 0, st, work, out, accumulate, nblk);
  return check_launch("reduce_final");
}

template <typename T>
int dispatch(int op, int B, int ny, int nx, const T* a, long long ap, long long ai, const T* b, long long bp,
             long long bi, double atol, double rtol, double* out, double* work, int acc, hipStream_t st) {
  switch (op) {
    case MJX_RED_SUMSQ: return launch<T, MJX_RED_SUMSQ>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    case MJX_RED_DOT: return launch<T, MJX_RED_DOT>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    case MJX_RED_MAXABS: return launch<T, MJX_RED_MAXABS>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    case MJX_RED_SCALED_SQ:
      return launch<T, MJX_RED_SCALED_SQ>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    case MJX_RED_NONFINITE:
      return launch<T, MJX_RED_NONFINITE>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    case MJX_RED_SUM: return launch<T, MJX_RED_SUM>(B, ny, nx, a, ap, ai, b, bp, bi, atol, rtol, out, work, acc, st);
    default: return fail(MJX_E_ARG, "unknown reduction op");
  }
}

}  // namespace
}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_reduce_workspace_elems(int B) { return B < 1 ? 0 : (long long)B * RED_BLOCKS; }

int mjx_reduce2d(int dtype, int op, int B, int ny, int nx, const void* a, long long a_pitch, long long a_istride,
                 const void* b, long long b_pitch, long long b_istride, double atol, double rtol, double* out,
                 int accumulate, double* work, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (B < 1 || ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty reduction");
  if (!a || !out || !work) return fail(MJX_E_ARG, "null pointer argument");
  if ((op == MJX_RED_DOT || op == MJX_RED_SCALED_SQ) && !b) return fail(MJX_E_ARG, "this reduction needs b");
  if (a_pitch < nx || (b && b_pitch < nx)) return fail(MJX_E_ARG, "pitch smaller than nx");
  if ((long long)B * RED_BLOCKS > 0x7fffffffLL) return fail(MJX_E_LIMIT, "batch too large for one launch");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MJX_F32)
    return dispatch<float>(op, B, ny, nx, (const float*)a, a_pitch, a_istride, (const float*)b, b_pitch, b_istride,
                           atol, rtol, out, work, accumulate, st);
  return dispatch<double>(op, B, ny, nx, (const double*)a, a_pitch, a_istride, (const double*)b, b_pitch,
                          b_istride, atol, rtol, out, work, accumulate, st);
}

int mjx_reduce_finish(int op, int B, const double* per_inst, int take_sqrt, double* total, void* stream) {
  if (B < 1 || !per_inst || !total) return fail(MJX_E_ARG, "empty batch or null pointer argument");
  if (op < MJX_RED_SUMSQ || op > MJX_RED_SUM) return fail(MJX_E_ARG, "unknown reduction op");
  hipStream_t st = (hipStream_t)stream;
  if (op == MJX_RED_MAXABS)
    hipLaunchKernelGGL((k_reduce_total<MJX_RED_MAXABS>), dim3(1), dim3(64), 0, st, per_inst, B, take_sqrt, total);
  else
    hipLaunchKernelGGL((k_reduce_total<MJX_RED_SUM>), dim3(1), dim3(64), 0, st, per_inst, B, take_sqrt, total);
  return check_launch("reduce_total");
}

}  // extern "C"
