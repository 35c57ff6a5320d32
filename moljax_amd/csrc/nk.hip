// Newton-Krylov for implicit reaction-diffusion steps: residual and
// hand-written Jacobian-vector product kernels, the Jacobi preconditioner
// sweep, deterministic FP64 Krylov vector kernels, and the native
// Newton / restarted GMRES driver.
//
// Problem class: unknown u on an interior (ny, nx) array, 5-point Laplacian
// with homogeneous Dirichlet boundaries on a vertex grid (zero outside) or
// periodic wrap, polynomial reaction f(u) = a3 u^3 + a2 u^2 + a1 u + a0:
//   R(u) = c u - u_old - b (D lap u + f(u)) - fold
//   J v  = c v - b (D lap v + f'(u) v)
// with (c, b, u_old, fold) = (1, theta dt, y_n, (1 - theta) dt F(y_n)) for
// backward Euler / Crank-Nicolson (moljax/core/newton_krylov.py:362-403) and
// (alpha0, beta, -(alpha1 y_n + alpha2 y_{n-1}), 0) for variable-step BDF2
// (newton_krylov.py:406-448).  The JVP replaces the forward-mode AD of
// newton_krylov.py:61-80.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <string>
#include <vector>

#include "czt_api.h"
#include "mjx_common.h"
#include "r2r_kernels.h"

namespace mjx {

struct RDArgs {
  int ny, nx, bc;  // bc 0: Dirichlet (vertex, zero outside), 1: periodic
  double dy2, dx2;
  double D, theta_dt, a3, a2, a1, a0;
  double lead;  // coefficient c of u in the residual (1 except BDF2)
};

template <typename T>
__device__ __forceinline__ T nb(const T* __restrict__ f, int y, int x, const RDArgs& a) {
  if (a.bc == 1) {
    y = y < 0 ? y + a.ny : (y >= a.ny ? y - a.ny : y);
    x = x < 0 ? x + a.nx : (x >= a.nx ? x - a.nx : x);
  } else if (y < 0 || y >= a.ny || x < 0 || x >= a.nx) {
    return (T)0;
  }
  return f[(long long)y * a.nx + x];
}

template <typename T>
__device__ __forceinline__ T lap5(const T* __restrict__ f, int y, int x, const RDArgs& a) {
  const T c = f[(long long)y * a.nx + x];
  const T yy = (nb(f, y + 1, x, a) - (T)2 * c + nb(f, y - 1, x, a)) / (T)a.dy2;
  const T xx = (nb(f, y, x + 1, a) - (T)2 * c + nb(f, y, x - 1, a)) / (T)a.dx2;
  return yy + xx;
}

// r = c u - u_old - b (D lap u + f(u)) - [fold term], FP64.
__global__ __launch_bounds__(256) void k_rd_residual(const double* __restrict__ u, const double* __restrict__ u_old,
                                                     const double* __restrict__ fold_term, double* __restrict__ r,
                                                     RDArgs a) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)a.ny * a.nx;
  if (g >= n) return;
  const int y = (int)(g / a.nx), x = (int)(g % a.nx);
  const double uu = u[g];
  const double f = ((a.a3 * uu + a.a2) * uu + a.a1) * uu + a.a0;
  double res = a.lead * uu - u_old[g] - a.theta_dt * (a.D * lap5(u, y, x, a) + f);
  if (fold_term) res -= fold_term[g];
  r[g] = res;
}

// w = c v - b (D lap v + f'(u) v); u in FP64, v / w in T.
template <typename T>
__global__ __launch_bounds__(256) void k_rd_jvp(const double* __restrict__ u, const T* __restrict__ v,
                                                T* __restrict__ w, RDArgs a) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)a.ny * a.nx;
  if (g >= n) return;
  const int y = (int)(g / a.nx), x = (int)(g % a.nx);
  const double uu = u[g];
  const T fp = (T)((3.0 * a.a3 * uu + 2.0 * a.a2) * uu + a.a1);
  const T vv = v[g];
  w[g] = (T)a.lead * vv - (T)a.theta_dt * ((T)a.D * lap5(v, y, x, a) + fp * vv);
}

// Weighted Jacobi sweep for (I - beta lap) x = r (moljax/core/preconditioners.py:166-198):
// x_new = omega (r + beta lap x) / diag + (1 - omega) x.
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi_sweep(const T* __restrict__ r, const T* __restrict__ x,
                                                      T* __restrict__ xn, RDArgs a, double beta, double omega,
                                                      double diag) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)a.ny * a.nx;
  if (g >= n) return;
  const int y = (int)(g / a.nx), x0 = (int)(g % a.nx);
  xn[g] = (T)omega * (r[g] + (T)beta * lap5(x, y, x0, a)) / (T)diag + ((T)1 - (T)omega) * x[g];
}

// ------------------------------------------------------ FP64 vector ops
// Wave64 shuffle reduction, then LDS across the block's waves.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int DOT_CHUNK = 8;    // basis vectors per block row
constexpr int DOT_BLOCKS = 256;  // fixed partition -> deterministic sums

// partial[(c * DOT_BLOCKS + blk) * DOT_CHUNK + i] = sum over this block's
// grid-stride slice of w * V[c*DOT_CHUNK + i]  (V rows of length n, stride n)
__global__ __launch_bounds__(256) void k_multidot_partial(const double* __restrict__ V, const double* __restrict__ w,
                                                          long long n, int k, double* __restrict__ partial) {
  const int c = blockIdx.y;
  const int i0 = c * DOT_CHUNK;
  double acc[DOT_CHUNK];
#pragma unroll
  for (int i = 0; i < DOT_CHUNK; ++i) acc[i] = 0.0;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)DOT_BLOCKS * blockDim.x) {
    const double wv = w[g];
#pragma unroll
    for (int i = 0; i < DOT_CHUNK; ++i)
      if (i0 + i < k) acc[i] += wv * V[(long long)(i0 + i) * n + g];
  }
  __shared__ double red[DOT_CHUNK][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < DOT_CHUNK; ++i) {
    const double s = wave_sum(acc[i]);
    if (lane == 0) red[i][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < DOT_CHUNK) {
    const int i = threadIdx.x;
    partial[((long long)c * DOT_BLOCKS + blockIdx.x) * DOT_CHUNK + i] = ((red[i][0] + red[i][1]) + red[i][2]) + red[i][3];
  }
}

// out[c*DOT_CHUNK + i] = sum_blk partial[...], fixed pairwise order.
__global__ __launch_bounds__(256) void k_multidot_final(const double* __restrict__ partial, int k,
                                                        double* __restrict__ out) {
  const int idx = blockIdx.x;  // basis index
  if (idx >= k) return;
  const int c = idx / DOT_CHUNK, i = idx % DOT_CHUNK;
  double v = partial[((long long)c * DOT_BLOCKS + threadIdx.x) * DOT_CHUNK + i];
  __shared__ double red[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[idx] = ((red[0] + red[1]) + red[2]) + red[3];
}

// w = beta_w * w + sum_i coef[i] * V[i]   (coef on the device)
__global__ __launch_bounds__(256) void k_multiaxpy(const double* __restrict__ V, const double* __restrict__ coef,
                                                   double sgn, int k, double* __restrict__ w, long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += coef[i] * V[(long long)i * n + g];
  w[g] += sgn * s;
}

// (x and y may alias)
__global__ __launch_bounds__(256) void k_scale_copy(const double* x, double a, double* y, long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) y[g] = a * x[g];
}

// (z may alias x or y)
__global__ __launch_bounds__(256) void k_axpby(double a, const double* x, double b, const double* y, double* z,
                                               long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) z[g] = a * x[g] + b * y[g];
}

__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ x, float* __restrict__ y, long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) y[g] = (float)x[g];
}

__global__ __launch_bounds__(256) void k_to_f64(const float* __restrict__ x, double* __restrict__ y, long long n) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) y[g] = (double)x[g];
}

static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }


// ------------------------------------------------------------- driver
struct NKConfig {
  int mv_dtype;        // MJX_F32 or MJX_F64: precision of the Krylov operator M A
  int precond;         // 0 none, 1 exact (I - dt D lap)^-1 (DST-I / FFT), 2 Jacobi
  double precond_dt;   // dt of the preconditioner context (preconditioners.py:29-43)
  int jacobi_iters;
  double jacobi_omega;
  int max_newton, max_krylov, restart, max_backtrack;
  double newton_tol, krylov_tol, damping, backtrack_factor, min_decrease;
};

struct NKStatsOut {
  int converged, newton_iters, gmres_iters, status;
  double res_norm;
};

class NKSolver {
 public:
  NKSolver(const RDArgs& a, const NKConfig& c, double* u, const double* u_old, const double* fold, char* work,
           const void* const* lam, hipStream_t st)
      : a_(a), c_(c), u_(u), u_old_(u_old), fold_(fold), st_(st) {
    n_ = (long long)a.ny * a.nx;
    const long long s2 = precond_work_reals(a.ny, a.nx);
    const int m = c.restart;
    double* p = (double*)work;
    V_ = p; p += (long long)(m + 1) * n_;
    w_ = p; p += n_;
    r_ = p; p += n_;
    dx_ = p; p += n_;
    xt_ = p; p += n_;
    t64_ = p; p += n_;
    t64b_ = p; p += n_;
    r2r64_ = p; p += s2;
    hdev_ = p; p += 3 * (m + 2);
    partial_ = p; p += (long long)((m + 1 + DOT_CHUNK - 1) / DOT_CHUNK) * DOT_BLOCKS * DOT_CHUNK;
    inst_ = p; p += 8;
    float* f = (float*)p;
    v32_ = f; f += n_;
    t32_ = f; f += n_;
    w32_ = f; f += n_;
    t32b_ = f; f += n_;
    r2r32_ = f; f += s2;
    lamx64_ = lam[0];
    lamy64_ = lam[1];
    lamx32_ = lam[2];
    lamy32_ = lam[3];
  }

  // Reals of the exact preconditioner's work array: the periodic spectral
  // work (complex, spec_work_cols per row), which also covers the n reals of
  // the DST solve.
  static long long precond_work_reals(int ny, int nx) { return 2LL * ny * spec_work_cols(nx, ny); }

  static long long workspace_bytes(int ny, int nx, int restart) {
    const long long n = (long long)ny * nx;
    const long long s2 = precond_work_reals(ny, nx);
    long long d = (long long)(restart + 1) * n + 6 * n + s2 + 3 * (restart + 2) +
                  (long long)((restart + 1 + DOT_CHUNK - 1) / DOT_CHUNK) * DOT_BLOCKS * DOT_CHUNK + 8;
    return d * 8 + (4 * n + s2) * 4 + 256;
  }

  int solve(NKStatsOut* out) {
    int rc = setup_inst();
    if (rc) return rc;
    double r0 = 0;
    if ((rc = residual(u_, r_, &r0))) return rc;
    bool converged = r0 < c_.newton_tol;
    double res_norm = r0;
    int it = 0, lin_total = 0;
    while (it < c_.max_newton && !converged) {
      double rn = 0;
      if ((rc = residual(u_, r_, &rn))) return rc;
      converged = rn < c_.newton_tol;
      // b = -r
      hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n_)), dim3(256), 0, st_, r_, -1.0, r_, n_);
      int lin = 0;
      if ((rc = gmres(r_, dx_, &lin))) return rc;
      lin_total += lin;
      // backtracking (newton_krylov.py:272-316): always max_backtrack trials
      double alpha = c_.damping, r_bt = rn;
      bool accepted = false;
      const double* x_bt = u_;
      for (int b = 0; b < c_.max_backtrack; ++b) {
        hipLaunchKernelGGL(k_axpby, dim3(nblk(n_)), dim3(256), 0, st_, 1.0, x_bt, alpha, (const double*)dx_, t64b_,
                           n_);
        double rnew = 0;
        if ((rc = residual(t64b_, t64_, &rnew))) return rc;
        const bool acc = rnew < (1.0 - c_.min_decrease * alpha) * r_bt;
        if (!(acc || accepted)) alpha *= c_.backtrack_factor;
        if (acc && !accepted) {
          copy(t64b_, xt_);
          x_bt = xt_;
          r_bt = rnew;
        }
        accepted = accepted || acc;
      }
      // fallback when the step is (numerically) not taken: x + damping dx,
      // as jnp.allclose(x_new, x) (rtol 1e-5, atol 1e-8) decides
      bool same = !accepted;
      if (accepted) {
        int sr = allclose(xt_, u_, &same);
        if (sr) return sr;
      }
      if (same) {
        hipLaunchKernelGGL(k_axpby, dim3(nblk(n_)), dim3(256), 0, st_, 1.0, (const double*)u_, c_.damping,
                           (const double*)dx_, u_, n_);
      } else {
        copy(xt_, u_);
      }
      res_norm = r_bt;
      ++it;
    }
    out->converged = converged ? 1 : 0;
    out->newton_iters = it;
    out->gmres_iters = lin_total;
    out->res_norm = res_norm;
    out->status = 0;
    return check_launch("newton");
  }

  // Left-preconditioned restarted GMRES(m) for J dx = b (b in bvec),
  // x0 = 0.  One iteration = one Arnoldi step; stops when the Givens
  // residual <= rtol * ||M b|| (checked every step); max_krylov restart cycles.
  int gmres(const double* bvec, double* x, int* iters_out) {
    const int m = c_.restart;
    int rc;
    hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n_)), dim3(256), 0, st_, bvec, 0.0, x, n_);
    if ((rc = precond(bvec, w_))) return rc;  // w = M b
    double bnorm = 0;
    if ((rc = norm(w_, &bnorm))) return rc;
    *iters_out = 0;
    if (bnorm == 0.0) return MJX_OK;
    const double target = c_.krylov_tol * bnorm;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), hh(3 * (m + 2));
    int iters = 0;
    double res = bnorm;
    for (int cyc = 0; cyc < c_.max_krylov; ++cyc) {
      if (iters > 0) {  // r = M (b - A x); xt_ is free during the Krylov solve
        if ((rc = apply_A(x, t64_))) return rc;
        hipLaunchKernelGGL(k_axpby, dim3(nblk(n_)), dim3(256), 0, st_, 1.0, bvec, -1.0, (const double*)t64_, xt_,
                           n_);
        if ((rc = precond(xt_, w_))) return rc;
      }
      double beta = 0;
      if ((rc = norm(w_, &beta))) return rc;
      if (beta <= target) break;
      hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)w_, 1.0 / beta, V_, n_);
      std::fill(g.begin(), g.end(), 0.0);
      std::fill(H.begin(), H.end(), 0.0);
      g[0] = beta;
      int k = 0;
      for (int j = 0; j < m; ++j) {
        if ((rc = apply_op(V_ + (long long)j * n_, w_))) return rc;
        ++iters;
        const int kk = j + 1;
        // CGS2: h = V^T w; w -= V h; h2 = V^T w; w -= V h2; hn = ||w||
        multidot(V_, w_, kk, hdev_);
        hipLaunchKernelGGL(k_multiaxpy, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)V_,
                           (const double*)hdev_, -1.0, kk, w_, n_);
        multidot(V_, w_, kk, hdev_ + (m + 2));
        hipLaunchKernelGGL(k_multiaxpy, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)V_,
                           (const double*)(hdev_ + (m + 2)), -1.0, kk, w_, n_);
        multidot(w_, w_, 1, hdev_ + 2 * (m + 2));
        if (hipMemcpyAsync(hh.data(), hdev_, sizeof(double) * 3 * (m + 2), hipMemcpyDeviceToHost, st_) != hipSuccess ||
            hipStreamSynchronize(st_) != hipSuccess)
          return fail(MJX_E_HIP, "gmres: device to host copy failed");
        const double hn = std::sqrt(hh[2 * (m + 2)]);
        for (int i = 0; i < kk; ++i) H[(size_t)i * m + j] = hh[i] + hh[(m + 2) + i];
        H[(size_t)(j + 1) * m + j] = hn;
        for (int i = 0; i < j; ++i) {
          const double t = cs[i] * H[(size_t)i * m + j] + sn[i] * H[(size_t)(i + 1) * m + j];
          H[(size_t)(i + 1) * m + j] = -sn[i] * H[(size_t)i * m + j] + cs[i] * H[(size_t)(i + 1) * m + j];
          H[(size_t)i * m + j] = t;
        }
        const double den = std::hypot(H[(size_t)j * m + j], H[(size_t)(j + 1) * m + j]);
        cs[j] = H[(size_t)j * m + j] / den;
        sn[j] = H[(size_t)(j + 1) * m + j] / den;
        H[(size_t)j * m + j] = den;
        H[(size_t)(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        res = std::fabs(g[j + 1]);
        k = kk;
        if (res <= target || hn == 0.0) break;
        hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)w_, 1.0 / hn,
                           V_ + (long long)kk * n_, n_);
      }
      for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int l = i + 1; l < k; ++l) s -= H[(size_t)i * m + l] * y[l];
        y[i] = s / H[(size_t)i * m + i];
      }
      if (hipMemcpyAsync(hdev_, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, st_) != hipSuccess)
        return fail(MJX_E_HIP, "gmres: host to device copy failed");
      hipLaunchKernelGGL(k_multiaxpy, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)V_, (const double*)hdev_, 1.0,
                         k, x, n_);
      if (hipStreamSynchronize(st_) != hipSuccess) return fail(MJX_E_HIP, "gmres: sync failed");
      if (res <= target) break;
    }
    *iters_out = iters;
    return check_launch("gmres");
  }

 private:
  int setup_inst() {
    const double row[8] = {c_.precond == 1 ? c_.precond_dt * a_.D : 0.0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(inst_, row, sizeof(row), hipMemcpyHostToDevice, st_) != hipSuccess)
      return fail(MJX_E_HIP, "nk: parameter upload failed");
    return hipStreamSynchronize(st_) == hipSuccess ? MJX_OK : fail(MJX_E_HIP, "nk: sync failed");
  }

  void copy(const double* x, double* y) {
    hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n_)), dim3(256), 0, st_, x, 1.0, y, n_);
  }

  void multidot(const double* V, const double* w, int k, double* out) {
    const int chunks = (k + DOT_CHUNK - 1) / DOT_CHUNK;
    hipLaunchKernelGGL(k_multidot_partial, dim3(DOT_BLOCKS, chunks), dim3(256), 0, st_, V, w, n_, k, partial_);
    hipLaunchKernelGGL(k_multidot_final, dim3(k), dim3(256), 0, st_, (const double*)partial_, k, out);
  }

  int norm(const double* x, double* out) {
    multidot(x, x, 1, hdev_ + 2 * (c_.restart + 2));
    double v = 0;
    if (hipMemcpyAsync(&v, hdev_ + 2 * (c_.restart + 2), sizeof(double), hipMemcpyDeviceToHost, st_) != hipSuccess ||
        hipStreamSynchronize(st_) != hipSuccess)
      return fail(MJX_E_HIP, "norm: copy failed");
    *out = std::sqrt(v);
    return MJX_OK;
  }

  int residual(const double* u, double* r, double* rnorm) {
    hipLaunchKernelGGL(k_rd_residual, dim3(nblk(n_)), dim3(256), 0, st_, u, u_old_, fold_, r, a_);
    return norm(r, rnorm);
  }

  int allclose(const double* a, const double* b, bool* same);

  // x (FP64) -> J x (FP64), always in FP64 (used for restart residuals).
  int apply_A(const double* x, double* y) {
    hipLaunchKernelGGL(k_rd_jvp<double>, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_, x, y, a_);
    return check_launch("jvp");
  }

  // w = M A v in the configured operator precision.
  int apply_op(const double* v, double* w) {
    if (c_.mv_dtype == MJX_F64) {
      hipLaunchKernelGGL(k_rd_jvp<double>, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_, v, t64_, a_);
      return precond(t64_, w);
    }
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(n_)), dim3(256), 0, st_, v, v32_, n_);
    hipLaunchKernelGGL(k_rd_jvp<float>, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_, (const float*)v32_,
                       t32_, a_);
    int rc = precond32(t32_, w32_);
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_f64, dim3(nblk(n_)), dim3(256), 0, st_, (const float*)w32_, w, n_);
    return check_launch("apply_op");
  }

  int precond(const double* in, double* out) {
    if (c_.mv_dtype == MJX_F32) {  // the whole operator M A stays in FP32
      hipLaunchKernelGGL(k_to_f32, dim3(nblk(n_)), dim3(256), 0, st_, in, v32_, n_);
      int rc = precond32(v32_, w32_);
      if (rc) return rc;
      hipLaunchKernelGGL(k_to_f64, dim3(nblk(n_)), dim3(256), 0, st_, (const float*)w32_, out, n_);
      return check_launch("precond");
    }
    return precond_t<double>(in, out, t64b_, r2r64_, lamx64_, lamy64_);
  }

  int precond32(const float* in, float* out) { return precond_t<float>(in, out, t32b_, r2r32_, lamx32_, lamy32_); }

  // in, out, tmp and work must be distinct buffers.
  template <typename T>
  int precond_t(const T* in, T* out, T* tmp, T* work, const void* lamx, const void* lamy) {
    const int dt = std::is_same<T, float>::value ? MJX_F32 : MJX_F64;
    if (c_.precond == 0) {
      return hipMemcpyAsync(out, in, sizeof(T) * n_, hipMemcpyDeviceToDevice, st_) == hipSuccess
                 ? MJX_OK
                 : fail(MJX_E_HIP, "precond copy failed");
    }
    if (c_.precond == 1) {
      if (a_.bc == 0)
        return mjx_r2r_solve2d(dt, 0, 1, a_.ny, a_.nx, in, a_.nx, n_, out, a_.nx, n_, lamx, lamy, inst_, 0, work,
                               st_);
      return mjx_spectral_pair(dt, 1, a_.ny, a_.nx, in, nullptr, a_.nx, n_, out, nullptr, a_.nx, n_, 0, lamx, lamy,
                               nullptr, inst_, 1, work, st_);
    }
    // Jacobi: x0 = r, n sweeps ping-ponging between out and tmp
    const double beta = c_.precond_dt * a_.D;
    const double diag = 1.0 + beta * (2.0 / a_.dx2 + 2.0 / a_.dy2);
    const T* x = in;
    for (int i = 0; i < c_.jacobi_iters; ++i) {
      T* dst = ((c_.jacobi_iters - 1 - i) % 2 == 0) ? out : tmp;
      hipLaunchKernelGGL(k_jacobi_sweep<T>, dim3(nblk(n_)), dim3(256), 0, st_, in, x, dst, a_, beta, c_.jacobi_omega,
                         diag);
      x = dst;
    }
    if (c_.jacobi_iters == 0)
      return hipMemcpyAsync(out, in, sizeof(T) * n_, hipMemcpyDeviceToDevice, st_) == hipSuccess
                 ? MJX_OK
                 : fail(MJX_E_HIP, "precond copy failed");
    return check_launch("jacobi");
  }

  RDArgs a_;
  NKConfig c_;
  double* u_;
  const double* u_old_;
  const double* fold_;
  hipStream_t st_;
  long long n_;
  double *V_, *w_, *r_, *dx_, *xt_, *t64_, *t64b_, *r2r64_, *hdev_, *partial_, *inst_;
  float *v32_, *t32_, *w32_, *t32b_, *r2r32_;
  const void *lamx64_, *lamy64_, *lamx32_, *lamy32_;
};

// flag = 1 if any |a - b| > atol + rtol |b| (jnp.allclose defaults).
__global__ __launch_bounds__(256) void k_allclose(const double* __restrict__ a, const double* __restrict__ b,
                                                  long long n, int* __restrict__ flag) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n && !(fabs(a[g] - b[g]) <= 1e-8 + 1e-5 * fabs(b[g]))) atomicOr(flag, 1);
}

int NKSolver::allclose(const double* a, const double* b, bool* same) {
  int* flag = (int*)(partial_);  // scratch: the dot partials are not live here
  int zero = 0, host = 0;
  if (hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, st_) != hipSuccess)
    return fail(MJX_E_HIP, "allclose: upload failed");
  hipLaunchKernelGGL(k_allclose, dim3(nblk(n_)), dim3(256), 0, st_, a, b, n_, flag);
  if (hipMemcpyAsync(&host, flag, sizeof(int), hipMemcpyDeviceToHost, st_) != hipSuccess ||
      hipStreamSynchronize(st_) != hipSuccess)
    return fail(MJX_E_HIP, "allclose: copy failed");
  *same = host == 0;
  return MJX_OK;
}

}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_nk_workspace_bytes(int ny, int nx, int restart) { return NKSolver::workspace_bytes(ny, nx, restart); }

int mjx_rd_newton_krylov(int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta,
                         const double* reaction, double* u, const double* u_old, const double* fold_term,
                         const void* const* lam_tables, const double* nk_params, const int* nk_iparams,
                         double* stats_out, void* work, void* stream) {
  if (ny < 2 || nx < 2 || (bc != 0 && bc != 1)) return fail(MJX_E_ARG, "bad grid or boundary kind");
  if (!reaction || !u || !u_old || !nk_params || !nk_iparams || !stats_out || !work || !lam_tables)
    return fail(MJX_E_ARG, "null pointer argument");
  RDArgs a;
  a.ny = ny;
  a.nx = nx;
  a.bc = bc;
  a.dy2 = dy * dy;
  a.dx2 = dx * dx;
  a.D = D;
  a.theta_dt = theta * dt;
  a.a3 = reaction[0];
  a.a2 = reaction[1];
  a.a1 = reaction[2];
  a.a0 = reaction[3];
  a.lead = nk_params[6];
  NKConfig c;
  c.mv_dtype = nk_iparams[0];
  c.precond = nk_iparams[1];
  c.jacobi_iters = nk_iparams[2];
  c.max_newton = nk_iparams[3];
  c.max_krylov = nk_iparams[4];
  c.restart = nk_iparams[5];
  c.max_backtrack = nk_iparams[6];
  c.newton_tol = nk_params[0];
  c.krylov_tol = nk_params[1];
  c.damping = nk_params[2];
  c.backtrack_factor = nk_params[3];
  c.min_decrease = nk_params[4];
  c.jacobi_omega = nk_params[5];
  c.precond_dt = nk_params[7];
  if (!(a.lead > 0)) return fail(MJX_E_ARG, "the residual's leading coefficient must be > 0");
  if (c.mv_dtype != MJX_F32 && c.mv_dtype != MJX_F64) return fail(MJX_E_ARG, "operator dtype must be F32 or F64");
  if (c.precond < 0 || c.precond > 2) return fail(MJX_E_ARG, "precond must be 0, 1 or 2");
  if (c.restart < 1 || c.restart > 200) return fail(MJX_E_LIMIT, "restart must be in [1, 200]");
  if (c.precond == 1 && (!lam_tables[0] || !lam_tables[1] || !lam_tables[2] || !lam_tables[3]))
    return fail(MJX_E_ARG, "exact preconditioner needs symbol tables");
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  NKSolver s(a, c, u, u_old, fold_term, (char*)work, lam_tables, (hipStream_t)stream);
  NKStatsOut out{};
  rc = s.solve(&out);
  stats_out[0] = out.converged;
  stats_out[1] = out.newton_iters;
  stats_out[2] = out.gmres_iters;
  stats_out[3] = out.res_norm;
  return rc;
}

int mjx_rd_residual(int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta, double lead,
                    const double* reaction, const double* u, const double* u_old, const double* fold_term, double* r,
                    void* stream) {
  if (!reaction || !u || !u_old || !r) return fail(MJX_E_ARG, "null pointer argument");
  RDArgs a{ny, nx, bc, dy * dy, dx * dx, D, theta * dt, reaction[0], reaction[1], reaction[2], reaction[3], lead};
  const long long n = (long long)ny * nx;
  hipLaunchKernelGGL(k_rd_residual, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, u, u_old, fold_term, r, a);
  return check_launch("rd_residual");
}

int mjx_rd_jvp(int dtype, int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta,
               double lead, const double* reaction, const double* u, const void* v, void* w, void* stream) {
  if (!reaction || !u || !v || !w) return fail(MJX_E_ARG, "null pointer argument");
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  RDArgs a{ny, nx, bc, dy * dy, dx * dx, D, theta * dt, reaction[0], reaction[1], reaction[2], reaction[3], lead};
  const long long n = (long long)ny * nx;
  if (dtype == MJX_F64)
    hipLaunchKernelGGL(k_rd_jvp<double>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, u, (const double*)v,
                       (double*)w, a);
  else
    hipLaunchKernelGGL(k_rd_jvp<float>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, u, (const float*)v,
                       (float*)w, a);
  return check_launch("rd_jvp");
}

int mjx_dot_f64(long long n, int k, const double* V, const double* w, double* out, double* partial, void* stream) {
  if (n < 1 || k < 1 || !V || !w || !out || !partial) return fail(MJX_E_ARG, "bad dot arguments");
  hipStream_t st = (hipStream_t)stream;
  const int chunks = (k + DOT_CHUNK - 1) / DOT_CHUNK;
  hipLaunchKernelGGL(k_multidot_partial, dim3(DOT_BLOCKS, chunks), dim3(256), 0, st, V, w, n, k, partial);
  hipLaunchKernelGGL(k_multidot_final, dim3(k), dim3(256), 0, st, (const double*)partial, k, out);
  return check_launch("dot_f64");
}

long long mjx_dot_partial_elems(int k) { return (long long)((k + DOT_CHUNK - 1) / DOT_CHUNK) * DOT_BLOCKS * DOT_CHUNK; }

}  // extern "C"
