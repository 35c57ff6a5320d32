// The explicit RK4 and the preconditioned Crank-Nicolson iteration of the
// Gray-Scott benchmark (benchmarks/benchmark_gray_scott.py:89-127 and
// :162-229), periodic, pseudo-spectral Laplacian, on the row / column
// spectral kernels.  The state stays physical between steps; every
// Laplacian is a row pass (x-FFT), a column pass (y-FFT, D lam, inverse
// y-FFT) and a row pass (inverse x-FFT) whose ends are fused with the
// method's pointwise work:
//
// RK4, per stage s = 1..4 (c = 0, dt/2, dt/2, dt):
//   pack    y_s = y + c k_{s-1}, x-FFT                                (row)
//   col     D lam                                                     (col)
//   unpack  inverse x-FFT -> D lap y_s; k_s = D lap y_s + R(y_s);
//           acc = k1, acc += 2 k2, acc += 2 k3, y += dt/6 (acc + k4)  (row)
// Each unpack but the call's last also forms the next stage's input (the
// next step's y after stage 4) and packs it, so a step is four row and four
// column passes.
//
// CN, per step: pack 2 y1 and set y0 <- y1, then up to max_iter iterations
//   col     D lam                                                     (col)
//   res     inverse x-FFT -> D lap(y1 + y0);
//           r = y1 - y0 - dt/2 (D lap(y1 + y0) + R(y1) + R(y0));
//           per-tile max |r_u|, max |r_v|; x-FFT of -r               (row)
//   norm    |r| = max |r_u| + max |r_v|; the next iteration runs while
//           |r| > tol and fewer than max_iter iterations ran           (tiny)
//   col     1 / (1 - dt/2 D lam)                                      (col)
//   update  inverse x-FFT -> d; y1 += d; when the next iteration runs,
//           x-FFT of y1 + y0                                          (row)
// Two masks per instance, in the aux columns of the instance rows: inst_lap
// (this iteration's residual, then the next iteration's) and inst_pc (this
// iteration's update).  Skipped instances leave every later kernel of the
// step untouched, so a step is a fixed launch sequence that a graph can
// hold, with the reference's data-dependent iteration count per instance.
#include <hip/hip_runtime.h>

#include <string>

#include "mjx_common.h"
#include "mjx_fft.h"
#include "spectral_api.h"
#include "spectral_kernels.h"

namespace mjx {
namespace {

enum { BOP_RK_PACK = 0, BOP_RK_UNPACK = 1, BOP_CN_PACK = 2, BOP_CN_RES = 3, BOP_CN_UPD = 4, BOP_RK_NEXT = 5 };

template <typename T>
struct BenchArgs {
  T* u;
  T* v;  // state (RK4: y; CN: y1), caller layout
  long long pitch, istride;
  T *p0, *p1, *p2, *p3;  // RK4: ku, kv, acc_u, acc_v; CN: u0, v0 (dense ny x nx planes)
  cpx<T>* W;
  const InstParams* ip;  // F, k, dt; aux = active mask when masked
  const InstParams* ip2;  // CN update: its aux is the update's mask
  double* partial;       // CN residual: (B, tiles, 2) max |r| per tile
  int ny, tiles, stage, masked;
};

template <typename T>
__device__ __forceinline__ T nanmax_t(T m, T x) {
  return (x > m || x != x) ? x : m;
}

template <typename T, int NX, int TR, int OP>
__global__ __launch_bounds__((RowCfg<NX, TR>::BT)) void k_row_bench(BenchArgs<T> a,
                                                                     const cpx<T>* __restrict__ tw) {
  using C = RowCfg<NX, TR>;
  __shared__ cpx<T> lds[C::TR * C::LP];
  __shared__ T red[2 * (C::BT / 64 + 1)];
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int inst = w / a.tiles;
  const int tile = w - inst * a.tiles;
  if (a.masked && (OP == BOP_CN_UPD ? a.ip2 : a.ip)[inst].aux <= 0.0) return;
  const int j = threadIdx.x % C::TN, rl = threadIdx.x / C::TN;
  const int y = tile * C::TR + rl;
  cpx<T>* lrow = lds + rl * C::LP;
  T* su = a.u + inst * a.istride + (long long)y * a.pitch;
  T* sv = a.v + inst * a.istride + (long long)y * a.pitch;
  const long long po = ((long long)inst * a.ny + y) * NX;
  T* q0 = a.p0 ? a.p0 + po : nullptr;
  T* q1 = a.p1 ? a.p1 + po : nullptr;
  T* q2 = a.p2 ? a.p2 + po : nullptr;
  T* q3 = a.p3 ? a.p3 + po : nullptr;
  cpx<T>* wrow = a.W + po;
  const T F = (T)a.ip[inst].F, kk = (T)a.ip[inst].k, dt = (T)a.ip[inst].dt;
  const T hdt = (T)(0.5 * a.ip[inst].dt);
  // RK4 stage offset c: 0, dt/2, dt/2, dt
  const T c = a.stage == 1 ? (T)0 : (a.stage == 4 ? dt : hdt);
  TwCache<T, NX> twc;
  load_twiddles(twc, j, tw);
  cpx<T> v[C::P];
  if constexpr (OP == BOP_RK_PACK || OP == BOP_CN_PACK) {
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const int x = j + C::TN * m;
      const T yu = su[x], yv = sv[x];
      if constexpr (OP == BOP_RK_PACK) {
        v[m] = a.stage == 1 ? mk(yu, yv) : mk(yu + c * q0[x], yv + c * q1[x]);
      } else {
        if (a.stage) {  // first iteration of the step: y0 <- y1
          q0[x] = yu;
          q1[x] = yv;
          v[m] = mk(yu + yu, yv + yv);
        } else {
          v[m] = mk(yu + q0[x], yv + q1[x]);
        }
      }
    }
    line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
  } else {
    // The pointwise operands, loaded before the transform so that their
    // latency overlaps the row's LDS exchanges.
    T yu_[C::P], yv_[C::P], a0_[C::P], a1_[C::P];
#pragma unroll
    for (int m = 0; m < C::P; ++m) {
      const int x = j + C::TN * m;
      yu_[m] = su[x];
      yv_[m] = sv[x];
      const bool need = OP == BOP_CN_RES || OP == BOP_CN_UPD || a.stage != 1;
      a0_[m] = need ? q0[x] : (T)0;
      a1_[m] = need ? q1[x] : (T)0;
    }
    load_unpack<T, NX, TR>(wrow, true, lrow, j, v);
    line_fft<T, NX, +1, 1, 1>(v, j, lrow, twc);
    if constexpr (OP == BOP_RK_UNPACK || OP == BOP_RK_NEXT) {
      const T w6 = (T)(a.ip[inst].dt / 6.0);
      const T cn = a.stage == 3 ? dt : hdt;  // the next stage's offset (stage 4 -> next step's y)
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int x = j + C::TN * m;
        const T yu = yu_[m], yv = yv_[m];
        T xu = yu, xv = yv;
        if (a.stage != 1) {
          xu = yu + c * a0_[m];
          xv = yv + c * a1_[m];
        }
        T ru, rv;
        gs_react(xu, xv, F, kk, ru, rv);
        const T ku = v[m].re + ru, kv = v[m].im + rv;
        if (a.stage == 1) {
          q2[x] = ku;
          q3[x] = kv;
        } else if (a.stage < 4) {
          q2[x] = q2[x] + (T)2 * ku;
          q3[x] = q3[x] + (T)2 * kv;
        } else {
          const T nu = yu + w6 * (q2[x] + ku), nv = yv + w6 * (q3[x] + kv);
          su[x] = nu;
          sv[x] = nv;
          if constexpr (OP == BOP_RK_NEXT) v[m] = mk(nu, nv);
        }
        if (a.stage < 4) {
          q0[x] = ku;
          q1[x] = kv;
          if constexpr (OP == BOP_RK_NEXT) v[m] = mk(yu + cn * ku, yv + cn * kv);
        }
      }
      if constexpr (OP == BOP_RK_UNPACK) return;
      line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
    } else if constexpr (OP == BOP_CN_UPD) {
      const bool next = a.ip[inst].aux > 0.0;
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const int x = j + C::TN * m;
        const T nu = yu_[m] + v[m].re, nv = yv_[m] + v[m].im;
        su[x] = nu;
        sv[x] = nv;
        if (next) v[m] = mk(nu + a0_[m], nv + a1_[m]);
      }
      if (!next) return;
      line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
    } else {  // BOP_CN_RES
      T mu = (T)0, mv = (T)0;
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        const T u1 = yu_[m], v1 = yv_[m], u0 = a0_[m], v0 = a1_[m];
        T r1u, r1v, r0u, r0v;
        gs_react(u1, v1, F, kk, r1u, r1v);
        gs_react(u0, v0, F, kk, r0u, r0v);
        const T resu = u1 - u0 - hdt * (v[m].re + r1u + r0u);
        const T resv = v1 - v0 - hdt * (v[m].im + r1v + r0v);
        mu = nanmax_t(mu, fabs(resu));
        mv = nanmax_t(mv, fabs(resv));
        v[m] = mk(-resu, -resv);
      }
      // workgroup max of (mu, mv): shuffles within each wave (or within the
      // whole workgroup when it is smaller than a wave), then across waves
      constexpr int WV = C::BT < 64 ? C::BT : 64;
#pragma unroll
      for (int off = WV / 2; off > 0; off >>= 1) {
        mu = nanmax_t(mu, __shfl_xor(mu, off, WV));
        mv = nanmax_t(mv, __shfl_xor(mv, off, WV));
      }
      const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
      if (lane == 0) {
        red[2 * wave] = mu;
        red[2 * wave + 1] = mv;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        T bu = red[0], bv = red[1];
        for (int q = 1; q < (C::BT + 63) / 64; ++q) {
          bu = nanmax_t(bu, red[2 * q]);
          bv = nanmax_t(bv, red[2 * q + 1]);
        }
        a.partial[2 * ((long long)inst * a.tiles + tile)] = (double)bu;
        a.partial[2 * ((long long)inst * a.tiles + tile) + 1] = (double)bv;
      }
      __syncthreads();
      line_fft<T, NX, -1, 1, 1>(v, j, lrow, twc);
    }
  }
#pragma unroll
  for (int m = 0; m < C::P; ++m) lwr(lrow, j + C::TN * m, v[m]);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < C::P; ++m) {
    const int p = j + C::TN * m;
    wrow[p] = pack_at<T, NX>(lrow, p);
  }
}

// Per-step reset of the CN masks and counters (one thread per instance).
__global__ void k_cn_begin(InstParams* lap, InstParams* pc, int* iters, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  lap[b].aux = 1.0;
  pc[b].aux = 1.0;
  iters[b] = 0;
}

// |r| = max |r_u| + max |r_v| over the instance's tiles (NaN propagates, and
// a NaN norm ends the iteration as the reference's `res_norm > tol` does).
// An instance that ran this iteration keeps its update (pc mask) and runs
// the next one while |r| > tol (lap mask); one that did not run neither.
// One wave per instance: each lane folds a strided share of the tiles, then
// the lanes combine by shuffles (the maximum does not depend on the order).
__global__ __launch_bounds__(64) void k_cn_norm(InstParams* lap, InstParams* pc, const double* partial, int tiles,
                                                int* iters, int* iters_total, double tol, int max_iter, int B) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  if (b >= B) return;
  if (lap[b].aux <= 0.0) {
    if (lane == 0) pc[b].aux = 0.0;
    return;
  }
  double mu = 0.0, mv = 0.0;
  for (int t = lane; t < tiles; t += 64) {
    mu = nanmax_t(mu, partial[2 * ((long long)b * tiles + t)]);
    mv = nanmax_t(mv, partial[2 * ((long long)b * tiles + t) + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mu = nanmax_t(mu, __shfl_xor(mu, off));
    mv = nanmax_t(mv, __shfl_xor(mv, off));
  }
  if (lane != 0) return;
  const double norm = mu + mv;
  const int it = iters[b] + 1;
  iters[b] = it;
  if (iters_total) iters_total[b] += 1;
  if (!(norm > tol) || it >= max_iter) lap[b].aux = 0.0;
}

template <typename T, int N, int OP>
struct BenchRow {
  template <int L>
  struct F {
    static int run(int B, BenchArgs<T> a, hipStream_t st) {
      a.tiles = a.ny / L;
      hipLaunchKernelGGL((k_row_bench<T, N, L, OP>), dim3(a.tiles * B), dim3(RowCfg<N, L>::BT), 0, st, a,
                         twiddles<T>());
      return check_launch("row_bench");
    }
  };
};

template <typename T, int OP>
int bench_row(int nx, int B, BenchArgs<T> a, hipStream_t st) {
  switch (nx) {
#define X(N) \
  case N:    \
    return with_lines<T, N, BenchRow<T, N, OP>::template F>(pick_lines(lines_for<T, N>(TK_ROW), a.ny), B, a, st);
    X(32) X(64) X(128) X(256) X(512) X(1024)
#undef X
    default:
      return fail(MJX_E_SIZE, "unsupported nx " + std::to_string(nx));
  }
}

template <typename T>
int bench_tiles(int nx, int ny) {
  switch (nx) {
#define X(N) \
  case N:    \
    return ny / pick_lines(lines_for<T, N>(TK_ROW), ny);
    X(32) X(64) X(128) X(256) X(512) X(1024)
#undef X
    default:
      return 0;
  }
}

struct BenchWork {
  void* W;
  void *p0, *p1, *p2, *p3;
  double* partial;
  int* iters;
};

BenchWork bench_planes(void* work, int es, int B, int ny, int nx) {
  const long long plane = (long long)B * ny * nx;
  char* p = (char*)work;
  BenchWork w;
  w.W = p;
  p += 2 * es * plane;
  w.p0 = p;
  w.p1 = p + es * plane;
  w.p2 = p + 2 * es * plane;
  w.p3 = p + 3 * es * plane;
  p += 4 * es * plane;
  w.partial = (double*)p;
  p += sizeof(double) * 2 * (long long)B * ny;
  w.iters = (int*)p;
  return w;
}

template <typename T>
int rk4_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, const T* lamx, const T* lamy,
            const InstParams* ip, int n_steps, void* work, hipStream_t st) {
  const BenchWork w = bench_planes(work, sizeof(T), B, ny, nx);
  BenchArgs<T> a{u, v, pitch, istride, (T*)w.p0, (T*)w.p1, (T*)w.p2, (T*)w.p3, (cpx<T>*)w.W, ip, nullptr, nullptr,
                 ny, 0, 1, 0};
  a.stage = 1;
  int rc = bench_row<T, BOP_RK_PACK>(nx, B, a, st);
  for (int s = 0; s < n_steps && !rc; ++s)
    for (int stage = 1; stage <= 4 && !rc; ++stage) {
      a.stage = stage;
      rc = spec_col_apply(ny, B, nx, (cpx<T>*)w.W, lamx, lamy, nullptr, ip, -1, st);
      const bool last = s == n_steps - 1 && stage == 4;
      if (!rc) rc = last ? bench_row<T, BOP_RK_UNPACK>(nx, B, a, st) : bench_row<T, BOP_RK_NEXT>(nx, B, a, st);
    }
  return rc;
}

template <typename T>
int cn_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, const T* lamx, const T* lamy,
           InstParams* lap, InstParams* pc, double tol, int max_iter, int n_steps, int* iters_total, void* work,
           hipStream_t st) {
  const BenchWork w = bench_planes(work, sizeof(T), B, ny, nx);
  BenchArgs<T> a{u, v, pitch, istride, (T*)w.p0, (T*)w.p1, nullptr, nullptr, (cpx<T>*)w.W, lap, pc, w.partial,
                 ny, 0, 0, 1};
  const int tiles = bench_tiles<T>(nx, ny);
  int rc = MJX_OK;
  for (int s = 0; s < n_steps && !rc; ++s) {
    hipLaunchKernelGGL(k_cn_begin, dim3((B + 63) / 64), dim3(64), 0, st, lap, pc, w.iters, B);
    rc = check_launch("cn_begin");
    a.stage = 1;
    if (!rc) rc = bench_row<T, BOP_CN_PACK>(nx, B, a, st);
    for (int it = 0; it < max_iter && !rc; ++it) {
      rc = spec_col_apply(ny, B, nx, (cpx<T>*)w.W, lamx, lamy, nullptr, lap, -1, st, MSTEP_ACTIVE);
      if (!rc) rc = bench_row<T, BOP_CN_RES>(nx, B, a, st);
      if (!rc) {
        hipLaunchKernelGGL(k_cn_norm, dim3(B), dim3(64), 0, st, lap, pc, w.partial, tiles, w.iters, iters_total,
                           tol, max_iter, B);
        rc = check_launch("cn_norm");
      }
      if (!rc) rc = spec_col_apply(ny, B, nx, (cpx<T>*)w.W, lamx, lamy, nullptr, pc, 1, st, MSTEP_ACTIVE);
      if (!rc) rc = bench_row<T, BOP_CN_UPD>(nx, B, a, st);
    }
  }
  return rc;
}

bool bench_size(int n) { return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024; }

int check_common(int dtype, int B, int ny, int nx, const void* u, const void* v, long long pitch, const void* lamx,
                 const void* lamy, const void* work, int n_steps) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (!bench_size(nx) || !bench_size(ny))
    return fail(MJX_E_SIZE, "grid " + std::to_string(ny) + "x" + std::to_string(nx) +
                                " not supported: ny, nx powers of two in [32, 1024]");
  if (B < 1 || !u || !v || !lamx || !lamy || !work) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  return MJX_OK;
}

}  // namespace
}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_gs_bench_workspace_bytes(int dtype, int B, int ny, int nx) {
  const long long es = dtype == MJX_F64 ? 8 : 4;
  const long long plane = (long long)B * ny * nx;
  return 6 * es * plane + 16LL * B * ny + 4LL * B;
}

int mjx_gs_rk4_spectral(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                        const void* lamx, const void* lamy, const double* inst, int n_steps, void* work,
                        void* stream) {
  int rc = check_common(dtype, B, ny, nx, u, v, pitch, lamx, lamy, work, n_steps);
  if (rc) return rc;
  if (!inst) return fail(MJX_E_ARG, "null inst");
  if (n_steps == 0) return MJX_OK;
  if ((rc = ensure_tables(nullptr))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (dtype == MJX_F32)
    return rk4_run<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, (const float*)lamx, (const float*)lamy,
                          ip, n_steps, work, st);
  return rk4_run<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, (const double*)lamx,
                         (const double*)lamy, ip, n_steps, work, st);
}

int mjx_gs_cn_fixed_point(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                          const void* lamx, const void* lamy, double* inst_lap, double* inst_pc, double tol,
                          int max_iter, int n_steps, int* iters_total, void* work, void* stream) {
  int rc = check_common(dtype, B, ny, nx, u, v, pitch, lamx, lamy, work, n_steps);
  if (rc) return rc;
  if (!inst_lap || !inst_pc) return fail(MJX_E_ARG, "null inst tables");
  if (max_iter < 1 || max_iter > 1000) return fail(MJX_E_ARG, "max_iter must be in [1, 1000]");
  if (n_steps == 0) return MJX_OK;
  if ((rc = ensure_tables(nullptr))) return rc;
  hipStream_t st = (hipStream_t)stream;
  InstParams* lap = reinterpret_cast<InstParams*>(inst_lap);
  InstParams* pc = reinterpret_cast<InstParams*>(inst_pc);
  if (dtype == MJX_F32)
    return cn_run<float>(B, ny, nx, (float*)u, (float*)v, pitch, istride, (const float*)lamx, (const float*)lamy, lap,
                         pc, tol, max_iter, n_steps, iters_total, work, st);
  return cn_run<double>(B, ny, nx, (double*)u, (double*)v, pitch, istride, (const double*)lamx, (const double*)lamy,
                        lap, pc, tol, max_iter, n_steps, iters_total, work, st);
}

}  // extern "C"
