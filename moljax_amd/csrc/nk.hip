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
#include <cstdint>
#include <mutex>
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

// ------------------------------------------------ device-resident GMRES(m)
// The Krylov loop keeps every scalar on the device: the Hessenberg column,
// the Givens rotations, the residual estimate and the convergence test are
// computed on the device, and later launches of a finished cycle exit at
// their first instruction.  The host enqueues Arnoldi steps in chunks and
// reads the control block back once per chunk (a few times per solve)
// instead of after every step.  Reductions end at kernel boundaries: RB
// workgroups write partials, and the next kernel finishes them in a fixed
// order, either in every one of its workgroups (the dot products that the
// basis update consumes, final_sums) or in one workgroup (the norm and the
// Givens step).  No cross-workgroup fence is needed and every sum has the
// same order on every run.
constexpr int RB = 256;  // workgroups of the partial reductions (fixed partition -> deterministic sums)
constexpr int RT = 256;  // threads per workgroup
constexpr int VC = 8;    // basis vectors per register chunk

struct GmresCtl {
  double beta, target, res, inv_hn;  // cycle norm, rtol * ||M b||, |g_{k}|, 1 / ||w|| of the last step
  double rtol, newton_tol, damping, bt_factor, min_dec;
  double rn, r_bt, alpha;  // Newton: residual norm at the iterate, backtracking state
  int k;                   // Arnoldi steps taken in the current cycle
  int iters;               // Arnoldi steps in this GMRES solve
  int done;                // cycle finished (converged or breakdown): later steps exit
  int conv;                // GMRES converged
  int newton_conv;         // rn < newton_tol at the current Newton iterate
  int accepted;            // a backtracking trial was accepted
  int moved;               // the accepted trial differs from the iterate (not allclose)
  int pad;
};

// Sum over the workgroup in a fixed tree (wave64 shuffles, then the waves in
// order); the result is valid in thread 0.
__device__ __forceinline__ double wg_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  return threadIdx.x == 0 ? ((red[0] + red[1]) + red[2]) + red[3] : 0.0;
}

// Per-workgroup dot products of w with basis vectors [0, k) over this
// workgroup's grid-stride slice, written as part[(c*RB + b)*VC + i].
__device__ void slice_dots(const double* __restrict__ V, const double* w, long long n, int k, double* part,
                           double* red) {
  const int nchunk = (k + VC - 1) / VC;
  for (int c = 0; c < nchunk; ++c) {
    double a[VC];
#pragma unroll
    for (int i = 0; i < VC; ++i) a[i] = 0.0;
    for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) {
      const double wv = w[g];
#pragma unroll
      for (int i = 0; i < VC; ++i)
        if (c * VC + i < k) a[i] += wv * V[(long long)(c * VC + i) * n + g];
    }
#pragma unroll
    for (int i = 0; i < VC; ++i) {
      const double s = wg_sum(a[i], red);
      if (threadIdx.x == 0) part[((long long)c * RB + blockIdx.x) * VC + i] = s;
    }
  }
}

// The fixed-order sums over workgroups of the partial columns idx < k
// (part[(c*RB + b)*VC + i], idx = c*VC + i), computed by every workgroup of
// the kernel that consumes them: thread t takes partial t of every index,
// then each index is summed over the wave (shuffle butterfly) and over the
// four waves in order, the tree of wg_sum, so every workgroup holds the same
// bits.  The sums land in hs (LDS); workgroup 0 also writes them to out.
__device__ void final_sums(const double* part, int k, double* hs, double* red, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c = 0; c * VC < k; ++c) {
    double v[VC];
#pragma unroll
    for (int i = 0; i < VC; ++i) v[i] = c * VC + i < k ? part[((long long)c * RB + threadIdx.x) * VC + i] : 0.0;
#pragma unroll
    for (int i = 0; i < VC; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < VC; ++i) red[i * 4 + wv] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < VC && c * VC + (int)threadIdx.x < k) {
      const int i = threadIdx.x;
      const double sum = ((red[i * 4] + red[i * 4 + 1]) + red[i * 4 + 2]) + red[i * 4 + 3];
      hs[c * VC + i] = sum;
      if (blockIdx.x == 0) out[c * VC + i] = sum;
    }
    __syncthreads();
  }
}

// Arrays of the Hessenberg / Givens recursion (m = restart length).
struct GivensArrays {
  const double* part;  // norm partials of the last CGS2 pass (RB values)
  const double* h1;    // CGS2 coefficients of the two passes (k values each)
  const double* h2;
  double* H;           // (m + 1) x m Hessenberg matrix, rotated
  double* cs;          // rotations
  double* sn;
  double* gv;          // rotated right-hand side (back substitution)
  double* ga;          // |g_j| before rotation j (ga[0] = beta)
};

// Givens step of Arnoldi step j: hn = ||w|| from the norm partials (the
// fixed-order tree of wg_sum), Hessenberg column j (h1 + h2, hn), the
// previous rotations, rotation j and the residual estimate |g_{j+1}| =
// |sn_j| |g_j|; the cycle ends on convergence (<= target) or breakdown
// (hn == 0).  Every workgroup that calls it computes the same bits from
// values written by earlier launches (h1, h2, cs/sn[0, j), ga[j]); only the
// writer stores column j, rotation j, g, ga[j + 1] and the control fields.
// Returns 1 / hn in *inv_hn (all threads) and whether the cycle is done.
__device__ bool givens_step(const GivensArrays& ga, int j, int m, GmresCtl* ctl, bool writer, double* red,
                            double* col, double* inv_hn) {
  __shared__ double bc[2];
  __shared__ double rc[2][256];  // rotations 0..j-1 (restart <= 200)
  const int k = j + 1;
  // the column and the rotations are loaded by the workgroup's threads in
  // parallel (the norm's barriers publish them), then thread 0 recurses
  const int t = threadIdx.x;
  if (t < k) col[t] = ga.h1[t] + ga.h2[t];
  if (t < j) {
    rc[0][t] = ga.cs[t];
    rc[1][t] = ga.sn[t];
  }
  const double hn2 = wg_sum(ga.part[t], red);
  if (t == 0) {
    const double hn = sqrt(hn2);
    col[j + 1] = hn;
    for (int i = 0; i < j; ++i) {
      const double a0 = col[i], a1 = col[i + 1];
      col[i] = rc[0][i] * a0 + rc[1][i] * a1;
      col[i + 1] = -rc[1][i] * a0 + rc[0][i] * a1;
    }
    const double hjj = col[j], hj1 = col[j + 1];
    const double den = hypot(hjj, hj1);
    const double c = hjj / den, s = hj1 / den;
    const double res = fabs(s) * ga.ga[j];
    const bool done = res <= ctl->target || hn == 0.0;
    if (writer) {
      for (int i = 0; i < j; ++i) ga.H[(long long)i * m + j] = col[i];
      ga.H[(long long)j * m + j] = den;
      ga.H[(long long)(j + 1) * m + j] = 0.0;
      ga.cs[j] = c;
      ga.sn[j] = s;
      const double gj = ga.gv[j];
      ga.gv[j + 1] = -s * gj;
      ga.gv[j] = c * gj;
      ga.ga[j + 1] = res;
      ctl->res = res;
      ctl->k = k;
      ctl->iters += 1;
      if (done) {
        ctl->done = 1;
        ctl->conv = res <= ctl->target ? 1 : 0;
      } else {
        ctl->inv_hn = 1.0 / hn;
      }
    }
    bc[0] = done ? 0.0 : 1.0 / hn;
    bc[1] = done ? 1.0 : 0.0;
  }
  __syncthreads();
  *inv_hn = bc[0];
  return bc[1] != 0.0;
}

// Start of a restart cycle from the norm partials of w (w = M (b - A x), or
// M b in cycle 0, where the target rtol * ||M b|| is set): beta = ||w|| and
// the cycle's initial state.  Computed the same way in every workgroup that
// calls it; only the writer stores it.  Returns 1 / beta in *inv_beta (all
// threads) and whether the cycle is done before any step.
__device__ bool cycle_start_step(const double* part, bool first, GmresCtl* ctl, double* gv, double* ga, bool writer,
                                 double* red, double* inv_beta) {
  __shared__ double bc[2];
  const double b2 = wg_sum(part[threadIdx.x], red);
  if (threadIdx.x == 0) {
    const double beta = sqrt(b2);
    const double target = first ? ctl->rtol * beta : ctl->target;
    const bool done = beta == 0.0 || beta <= target;
    if (writer) {
      ctl->k = 0;
      if (first) {
        ctl->iters = 0;
        ctl->target = target;
      }
      ctl->beta = beta;
      if (done) {
        ctl->done = 1;
        ctl->conv = 1;
      } else {
        ctl->done = 0;
        ctl->conv = 0;
        gv[0] = beta;
        ga[0] = beta;
        ctl->inv_hn = 1.0 / beta;
      }
    }
    bc[0] = done ? 0.0 : 1.0 / beta;
    bc[1] = done ? 1.0 : 0.0;
  }
  __syncthreads();
  *inv_beta = bc[0];
  return bc[1] != 0.0;
}

// Step j, part 1: v_j = w * inv_hn (written to V_j), t = J v_j in T.  With
// JVP_GIVENS the kernel first makes the Givens step of Arnoldi step j - 1,
// with JVP_START (step 0 of the first cycle) the cycle start, from the norm
// partials in gva.part (in every workgroup, workgroup 0 storing the
// result), which saves a one-workgroup launch; with JVP_PLAIN 1 / hn comes
// from the control block.
enum { JVP_PLAIN = 0, JVP_GIVENS = 1, JVP_START = 2 };

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_gm_jvp(const double* __restrict__ u, const double* __restrict__ w,
                                                GmresCtl* ctl, double* __restrict__ Vj, T* __restrict__ t,
                                                RDArgs a, GivensArrays gva, int jprev, int m) {
  double s;
  if constexpr (MODE == JVP_START) {
    // a cycle start resets the done flag of the previous solve: no check
    __shared__ double red[4];
    if (cycle_start_step(gva.part, true, ctl, gva.gv, gva.ga, blockIdx.x == 0, red, &s)) return;
  } else if constexpr (MODE == JVP_GIVENS) {
    // the control block's done flag as of this launch, read once per
    // workgroup (workgroup 0 may set it during the launch)
    __shared__ int done0;
    __shared__ double red[4];
    __shared__ double col[256];  // restart <= 200
    if (threadIdx.x == 0) done0 = ctl->done;
    __syncthreads();
    if (done0) return;
    if (givens_step(gva, jprev, m, ctl, blockIdx.x == 0, red, col, &s)) return;
  } else {
    if (ctl->done) return;
    s = ctl->inv_hn;
  }
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)a.ny * a.nx;
  if (g >= n) return;
  const int y = (int)(g / a.nx), x = (int)(g % a.nx);
  const double vc = w[g] * s;
  Vj[g] = vc;
  auto vat = [&](int yy, int xx) -> T {
    if (a.bc == 1) {
      yy = yy < 0 ? yy + a.ny : (yy >= a.ny ? yy - a.ny : yy);
      xx = xx < 0 ? xx + a.nx : (xx >= a.nx ? xx - a.nx : xx);
    } else if (yy < 0 || yy >= a.ny || xx < 0 || xx >= a.nx) {
      return (T)0;
    }
    return (T)(w[(long long)yy * a.nx + xx] * s);
  };
  const T c = (T)vc;
  const T lap = (vat(y + 1, x) - (T)2 * c + vat(y - 1, x)) / (T)a.dy2 +
                (vat(y, x + 1) - (T)2 * c + vat(y, x - 1)) / (T)a.dx2;
  const double uu = u[g];
  const T fp = (T)((3.0 * a.a3 * uu + 2.0 * a.a2) * uu + a.a1);
  t[g] = (T)a.lead * c - (T)a.theta_dt * ((T)a.D * lap + fp * c);
}

// Step j, part 2: partials of h1 = V[0..k)^T w (w given in TIN; for float it
// is widened into w64 on the way).
template <typename TIN>
__global__ __launch_bounds__(256) void k_gm_dot1(const double* __restrict__ V, const TIN* win, double* w64,
                                                 long long n, int k, double* part, const GmresCtl* ctl) {
  if (ctl->done) return;
  __shared__ double red[4];
  if constexpr (!std::is_same<TIN, double>::value) {
    for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT)
      w64[g] = (double)win[g];
  }
  slice_dots(V, w64, n, k, part, red);
}

// Step j, part 3: h1 from the partials of part 2 (also written to h1),
// w -= V h1, then partials of h2 = V^T w.
__global__ __launch_bounds__(256) void k_gm_upd_dot2(const double* __restrict__ V, double* w, long long n, int k,
                                                     const double* __restrict__ part_in, double* __restrict__ h1,
                                                     double* __restrict__ part, const GmresCtl* ctl) {
  if (ctl->done) return;
  __shared__ double red[4];
  __shared__ double redk[4 * VC];
  __shared__ double hs[256];
  final_sums(part_in, k, hs, redk, h1);
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += hs[i] * V[(long long)i * n + g];
    w[g] -= s;
  }
  slice_dots(V, w, n, k, part, red);
}

// Step j, part 4: h2 from the partials of part 3 (also written to h2),
// w -= V h2 and partials of ||w||^2.
__global__ __launch_bounds__(256) void k_gm_upd_norm(const double* __restrict__ V, double* w, long long n, int k,
                                                     const double* __restrict__ part_in, double* __restrict__ h2,
                                                     double* __restrict__ part, const GmresCtl* ctl) {
  if (ctl->done) return;
  __shared__ double red[4];
  __shared__ double redk[4 * VC];
  __shared__ double hs[256];
  final_sums(part_in, k, hs, redk, h2);
  double acc = 0.0;
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += hs[i] * V[(long long)i * n + g];
    const double wn = w[g] - s;
    w[g] = wn;
    acc += wn * wn;
  }
  const double ps = wg_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = ps;
}

// Step j, part 5 (one workgroup): the Givens step of Arnoldi step j, where
// the next step's JVP does not make it (the last step of a chunk).
__global__ __launch_bounds__(256) void k_gm_givens(GivensArrays gva, int j, int m, GmresCtl* ctl) {
  if (ctl->done) return;
  __shared__ double red[4];
  __shared__ double col[256];  // restart <= 200
  double inv_hn;
  givens_step(gva, j, m, ctl, true, red, col, &inv_hn);
}

// Partials of ||w||^2 (no control check: cycle starts always run).
__global__ __launch_bounds__(256) void k_norm2_partial(const double* w, long long n, double* part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) acc += w[g] * w[g];
  const double ps = wg_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = ps;
}

// Cycle start in one workgroup (restart cycles, whose start the host reads).
__global__ __launch_bounds__(256) void k_gm_cycle_start(const double* part, int first, GmresCtl* ctl, double* gv,
                                                        double* ga) {
  __shared__ double red[4];
  double inv_beta;
  cycle_start_step(part, first != 0, ctl, gv, ga, true, red, &inv_beta);
}

// End of a cycle: y = H[:k,:k]^-1 g[:k] (back substitution, repeated in every
// workgroup) and x += V[:k]^T y.
__global__ __launch_bounds__(256) void k_gm_update_x(const double* __restrict__ V, long long n, int m,
                                                     const double* __restrict__ H, const double* __restrict__ gv,
                                                     const GmresCtl* ctl, double* __restrict__ x) {
  const int k = ctl->k;
  if (k == 0) return;
  __shared__ double y[256];
  if (threadIdx.x == 0) {
    for (int i = k - 1; i >= 0; --i) {
      double s = gv[i];
      for (int l = i + 1; l < k; ++l) s -= H[(long long)i * m + l] * y[l];
      y[i] = s / H[(long long)i * m + i];
    }
  }
  __syncthreads();
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)gridDim.x * RT) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += y[i] * V[(long long)i * n + g];
    x[g] += s;
  }
}

// ------------------------------------------------ device-resident Newton
// Residual at the iterate: b = -R(u) and partials of ||R(u)||^2.
__global__ __launch_bounds__(256) void k_nk_residual(const double* __restrict__ u, const double* __restrict__ u_old,
                                                     const double* __restrict__ fold, double* __restrict__ b,
                                                     RDArgs a, double* part) {
  __shared__ double red[4];
  const long long n = (long long)a.ny * a.nx;
  double acc = 0.0;
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) {
    const int y = (int)(g / a.nx), x = (int)(g % a.nx);
    const double uu = u[g];
    const double f = ((a.a3 * uu + a.a2) * uu + a.a1) * uu + a.a0;
    double r = a.lead * uu - u_old[g] - a.theta_dt * (a.D * lap5(u, y, x, a) + f);
    if (fold) r -= fold[g];
    b[g] = -r;
    acc += r * r;
  }
  const double ps = wg_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = ps;
}

// rn = ||R(u)|| from the residual partials, the convergence flag, and the
// backtracking state of the iteration reset (stored by the writer; every
// thread of the workgroup calls it).
__device__ void residual_finish(const double* part, GmresCtl* ctl, bool writer, double* red) {
  const double r2 = wg_sum(part[threadIdx.x], red);
  if (threadIdx.x != 0 || !writer) return;
  const double rn = sqrt(r2);
  ctl->rn = rn;
  ctl->newton_conv = rn < ctl->newton_tol ? 1 : 0;
  ctl->r_bt = rn;
  ctl->alpha = ctl->damping;
  ctl->accepted = 0;
  ctl->moved = 0;
}

// One workgroup: the residual finish (before the host reads it).
__global__ __launch_bounds__(256) void k_nk_residual_final(const double* part, GmresCtl* ctl) {
  __shared__ double red[4];
  residual_finish(part, ctl, true, red);
}

// GMRES start x = 0 b (as a scale of b, so non-finite entries propagate);
// with RES workgroup 0 also finishes the Newton residual from its partials,
// which saves that one-workgroup launch.
template <bool RES>
__global__ __launch_bounds__(256) void k_gm_zero_x(const double* b, double* x, long long n, const double* part,
                                                   GmresCtl* ctl) {
  if constexpr (RES) {
    __shared__ double red[4];
    if (blockIdx.x == 0) residual_finish(part, ctl, true, red);
  }
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) x[g] = 0.0 * b[g];
}

// Acceptance of a backtracking trial from its residual partials: accept
// when ||R(x_t)|| < (1 - min_dec alpha) r_bt, r_bt being the iterate's
// residual norm rn until a trial is accepted (no trial runs after that);
// otherwise the next trial halves alpha.  Computed the same way in every
// workgroup that calls it; only the writer stores the result.
__device__ bool trial_accepts(const double* part, double alpha, GmresCtl* ctl, bool writer, double* red) {
  __shared__ int acc;
  const double r2 = wg_sum(part[threadIdx.x], red);
  if (threadIdx.x == 0) {
    const double rnew = sqrt(r2);
    const bool a = rnew < (1.0 - ctl->min_dec * alpha) * ctl->rn;
    if (writer) {
      if (a) {
        ctl->accepted = 1;
        ctl->r_bt = rnew;
      } else {
        ctl->alpha = alpha * ctl->bt_factor;
      }
    }
    acc = a ? 1 : 0;
  }
  __syncthreads();
  return acc != 0;
}

// alpha of trial t: damping halved t times, in the order of the sequential
// updates
__device__ __forceinline__ double trial_alpha(const GmresCtl* ctl, int t) {
  double a = ctl->damping;
  for (int i = 0; i < t; ++i) a = a * ctl->bt_factor;
  return a;
}

// The accepted flag as of this launch, read once per workgroup (workgroup 0
// of a launch that decides the previous trial may set it meanwhile).
__device__ __forceinline__ bool accepted_at_launch(const GmresCtl* ctl) {
  __shared__ int a0;
  if (threadIdx.x == 0) a0 = ctl->accepted;
  __syncthreads();
  return a0 != 0;
}

// Backtracking trial t (newton_krylov.py:272-316): x_t = u + alpha_t dx and
// partials of ||R(x_t)||^2 (into part_out), its neighbours formed the same
// way.  With FOLD the kernel first decides trial t - 1 from its partials
// (part_prev) in every workgroup, workgroup 0 storing the decision, which
// saves that decision's one-workgroup launch.  Trials after an acceptance
// leave the state unchanged in the reference, so they are skipped.
template <bool FOLD>
__global__ __launch_bounds__(256) void k_nk_trial(const double* __restrict__ u, const double* __restrict__ dx,
                                                  double* __restrict__ xt, const double* __restrict__ u_old,
                                                  const double* __restrict__ fold, RDArgs a, const double* part_prev,
                                                  double* part_out, int t, GmresCtl* ctl) {
  if (accepted_at_launch(ctl)) return;
  __shared__ double red[4];
  if constexpr (FOLD) {
    if (trial_accepts(part_prev, trial_alpha(ctl, t - 1), ctl, blockIdx.x == 0, red)) return;
  }
  const double alpha = trial_alpha(ctl, t);
  const long long n = (long long)a.ny * a.nx;
  auto xa = [&](int yy, int xx) -> double {
    if (a.bc == 1) {
      yy = yy < 0 ? yy + a.ny : (yy >= a.ny ? yy - a.ny : yy);
      xx = xx < 0 ? xx + a.nx : (xx >= a.nx ? xx - a.nx : xx);
    } else if (yy < 0 || yy >= a.ny || xx < 0 || xx >= a.nx) {
      return 0.0;
    }
    const long long q = (long long)yy * a.nx + xx;
    const double ad = alpha * dx[q];
    return u[q] + ad;
  };
  double acc = 0.0;
  for (long long g = (long long)blockIdx.x * RT + threadIdx.x; g < n; g += (long long)RB * RT) {
    const int y = (int)(g / a.nx), x = (int)(g % a.nx);
    const double c = xa(y, x);
    xt[g] = c;
    const double lap = (xa(y + 1, x) - 2.0 * c + xa(y - 1, x)) / a.dy2 + (xa(y, x + 1) - 2.0 * c + xa(y, x - 1)) / a.dx2;
    const double f = ((a.a3 * c + a.a2) * c + a.a1) * c + a.a0;
    double r = a.lead * c - u_old[g] - a.theta_dt * (a.D * lap + f);
    if (fold) r -= fold[g];
    acc += r * r;
  }
  const double ps = wg_sum(acc, red);
  if (threadIdx.x == 0) part_out[blockIdx.x] = ps;
}

// moved = 1 if any |x_t - u| > 1e-8 + 1e-5 |u| (jnp.allclose defaults), for
// an accepted trial.  With FOLD the kernel first decides the last trial (t)
// from its partials, as k_nk_trial does for the one before it.
template <bool FOLD>
__global__ __launch_bounds__(256) void k_nk_allclose(const double* __restrict__ xt, const double* __restrict__ u,
                                                     long long n, const double* part_last, int t, GmresCtl* ctl) {
  bool acc = accepted_at_launch(ctl);
  if constexpr (FOLD) {
    __shared__ double red[4];
    if (!acc) acc = trial_accepts(part_last, trial_alpha(ctl, t), ctl, blockIdx.x == 0, red);
  }
  if (!acc) return;
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n && !(fabs(xt[g] - u[g]) <= 1e-8 + 1e-5 * fabs(u[g]))) any = 1;
  __syncthreads();
  if (threadIdx.x == 0 && any) atomicOr(&ctl->moved, 1);  // one atomic per workgroup
}

// The Newton update: the accepted trial, or u + damping dx when no trial was
// accepted or the accepted one did not move the iterate.
__global__ __launch_bounds__(256) void k_nk_update(double* __restrict__ u, const double* __restrict__ xt,
                                                   const double* __restrict__ dx, long long n, const GmresCtl* ctl) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  if (ctl->accepted && ctl->moved) {
    u[g] = xt[g];
  } else {
    const double ad = ctl->damping * dx[g];
    u[g] = u[g] + ad;
  }
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

// Pinned host copy of the control block, one per device (read back once per
// chunk of Arnoldi steps).
static std::mutex g_ctl_mu;
static GmresCtl* g_ctl_host[64] = {nullptr};

static GmresCtl* ctl_host() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_ctl_mu);
  if (!g_ctl_host[dev]) {
    void* p = nullptr;
    if (hipHostMalloc(&p, sizeof(GmresCtl), hipHostMallocDefault) != hipSuccess) return nullptr;
    g_ctl_host[dev] = (GmresCtl*)p;
  }
  return g_ctl_host[dev];
}

// Arnoldi steps enqueued before the first read-back of a solve: the count
// of the previous solve at the same Newton iteration (successive implicit
// steps repeat their per-iteration counts closely), a restart cycle at most.
constexpr int NK_HINTS = 16;
static thread_local int g_gmres_hint[NK_HINTS] = {4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

class NKSolver {
 public:
  NKSolver(const RDArgs& a, const NKConfig& c, double* u, const double* u_old, const double* fold, char* work,
           const void* const* lam, hipStream_t st)
      : a_(a), c_(c), u_(u), u_old_(u_old), fold_(fold), st_(st) {
    n_ = (long long)a.ny * a.nx;
    const int m = c.restart;
    char* base = work;
    ctl_ = (GmresCtl*)base;
    double* p = (double*)(base + 256);
    H_ = p; p += (long long)(m + 1) * m;
    cs_ = p; p += m;
    sn_ = p; p += m;
    g_ = p; p += m + 1;
    h1_ = p; p += m + 1;
    h2_ = p; p += m + 1;
    ga_ = p; p += m + 1;
    part_ = p; p += part_elems(m);
    part2_ = p; p += part_elems(m);
    p = (double*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    V_ = p; p += (long long)(m + 1) * n_;
    w_ = p; p += n_;
    b_ = p; p += n_;
    dx_ = p; p += n_;
    xt_ = p; p += n_;
    t64_ = p; p += n_;
    t64b_ = p; p += n_;
    const long long s2 = precond_work_reals(a.ny, a.nx);
    r2r64_ = p; p += s2;
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

  static long long part_elems(int m) { return (long long)((m + 1 + VC - 1) / VC) * RB * VC; }

  // Reals of the exact preconditioner's work array: the periodic spectral
  // work (complex, spec_work_cols per row), which also covers the n reals of
  // the DST solve.
  static long long precond_work_reals(int ny, int nx) { return 2LL * ny * spec_work_cols(nx, ny); }

  static long long workspace_bytes(int ny, int nx, int restart) {
    const long long n = (long long)ny * nx;
    const int m = restart;
    const long long s2 = precond_work_reals(ny, nx);
    const long long small = (long long)(m + 1) * m + 2LL * m + 4LL * (m + 1) + 2 * part_elems(m);
    long long d = (long long)(m + 1) * n + 6 * n + s2 + 8;
    return 256 + small * 8 + 256 + d * 8 + (4 * n + s2) * 4 + 256;
  }

  int solve(NKStatsOut* out) {
    int rc = setup();
    if (rc) return rc;
    GmresCtl* hc = ctl_host();
    if (!hc) return fail(MJX_E_HIP, "nk: pinned control block unavailable");
    // residual at the initial guess; its flags serve the first iteration
    residual(true);
    if ((rc = read_ctl(hc))) return rc;
    bool converged = hc->newton_conv != 0;
    int it = 0, lin_total = 0;
    while (it < c_.max_newton && !converged) {
      if (it > 0) residual(false);
      int lin = 0;
      if ((rc = gmres(b_, dx_, hc, &lin, std::min(it, NK_HINTS - 1), it > 0))) return rc;
      lin_total += lin;
      converged = hc->newton_conv != 0;  // this iteration's test (made before its update)
      // trial t writes its partials to buffer t % 2; the next trial (or the
      // allclose pass after the last) decides it
      double* tp[2] = {part_, part2_};
      const int nbt = c_.max_backtrack;
      for (int t = 0; t < nbt; ++t) {
        if (t == 0)
          hipLaunchKernelGGL((k_nk_trial<false>), dim3(RB), dim3(RT), 0, st_, (const double*)u_, (const double*)dx_,
                             xt_, u_old_, fold_, a_, (const double*)nullptr, tp[0], 0, ctl_);
        else
          hipLaunchKernelGGL((k_nk_trial<true>), dim3(RB), dim3(RT), 0, st_, (const double*)u_, (const double*)dx_,
                             xt_, u_old_, fold_, a_, (const double*)tp[(t - 1) % 2], tp[t % 2], t, ctl_);
      }
      if (nbt > 0)
        hipLaunchKernelGGL((k_nk_allclose<true>), dim3(nblk(n_)), dim3(256), 0, st_, (const double*)xt_,
                           (const double*)u_, n_, (const double*)tp[(nbt - 1) % 2], nbt - 1, ctl_);
      else
        hipLaunchKernelGGL((k_nk_allclose<false>), dim3(nblk(n_)), dim3(256), 0, st_, (const double*)xt_,
                           (const double*)u_, n_, (const double*)nullptr, 0, ctl_);
      hipLaunchKernelGGL(k_nk_update, dim3(nblk(n_)), dim3(256), 0, st_, u_, (const double*)xt_,
                         (const double*)dx_, n_, (const GmresCtl*)ctl_);
      ++it;
    }
    if ((rc = read_ctl(hc))) return rc;
    out->converged = converged ? 1 : 0;
    out->newton_iters = it;
    out->gmres_iters = lin_total;
    out->res_norm = it > 0 ? hc->r_bt : hc->rn;
    out->status = 0;
    return check_launch("newton");
  }

  // Left-preconditioned restarted GMRES(m) for J dx = b, x0 = 0.  One
  // iteration = one Arnoldi step; the Givens residual is compared with rtol *
  // ||M b|| after every step; at most max_krylov restart cycles.
  // res_pending: the Newton residual's partials in part_ still need their
  // finish (made by the first launch here)
  int gmres(const double* bvec, double* x, GmresCtl* hc, int* iters_out, int hint_slot, bool res_pending) {
    const int m = c_.restart;
    int rc;
    if (res_pending)
      hipLaunchKernelGGL((k_gm_zero_x<true>), dim3(nblk(n_)), dim3(256), 0, st_, bvec, x, n_, (const double*)part_,
                         ctl_);
    else
      hipLaunchKernelGGL((k_gm_zero_x<false>), dim3(nblk(n_)), dim3(256), 0, st_, bvec, x, n_,
                         (const double*)nullptr, ctl_);
    if ((rc = precond(bvec, w_))) return rc;  // w = M b
    int iters = 0;
    for (int cyc = 0; cyc < c_.max_krylov; ++cyc) {
      if (cyc > 0) {  // w = M (b - A x)
        if ((rc = apply_A(x, t64_))) return rc;
        hipLaunchKernelGGL(k_axpby, dim3(nblk(n_)), dim3(256), 0, st_, 1.0, bvec, -1.0, (const double*)t64_, xt_,
                           n_);
        if ((rc = precond(xt_, w_))) return rc;
      }
      hipLaunchKernelGGL(k_norm2_partial, dim3(RB), dim3(RT), 0, st_, (const double*)w_, n_, part_);
      if (cyc > 0) {  // a restart may find the target met before any step
        hipLaunchKernelGGL(k_gm_cycle_start, dim3(1), dim3(RT), 0, st_, (const double*)part_, 0, ctl_, g_, ga_);
        if ((rc = read_ctl(hc))) return rc;
        if (hc->done) break;
      }
      int j = 0;
      int chunk = std::max(1, std::min(m, g_gmres_hint[hint_slot]));
      while (j < m) {
        const int c = std::min(chunk, m - j);
        // each step's JVP makes the previous step's Givens step, except the
        // chunk's first; the chunk's last step launches its own before the
        // control block is read
        // (the first cycle's first step also makes the cycle start)
        for (int jj = j; jj < j + c; ++jj) {
          const int mode = jj > j ? JVP_GIVENS : (cyc == 0 && jj == 0 ? JVP_START : JVP_PLAIN);
          if ((rc = arnoldi_step(jj, mode, jj == j + c - 1))) return rc;
        }
        j += c;
        if ((rc = read_ctl(hc))) return rc;
        if (hc->done) break;
        chunk = 2;
      }
      iters = hc->iters;
      hipLaunchKernelGGL(k_gm_update_x, dim3(RB), dim3(RT), 0, st_, (const double*)V_, n_, m, (const double*)H_,
                         (const double*)g_, (const GmresCtl*)ctl_, x);
      if (hc->conv) break;
    }
    iters = hc->iters;
    if (iters > 0) g_gmres_hint[hint_slot] = iters;
    *iters_out = iters;
    return check_launch("gmres");
  }

 private:
  int setup() {
    GmresCtl c0{};
    c0.rtol = c_.krylov_tol;
    c0.newton_tol = c_.newton_tol;
    c0.damping = c_.damping;
    c0.bt_factor = c_.backtrack_factor;
    c0.min_dec = c_.min_decrease;
    c0.done = 1;
    const double row[8] = {c_.precond == 1 ? c_.precond_dt * a_.D : 0.0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(ctl_, &c0, sizeof(c0), hipMemcpyHostToDevice, st_) != hipSuccess ||
        hipMemcpyAsync(inst_, row, sizeof(row), hipMemcpyHostToDevice, st_) != hipSuccess)
      return fail(MJX_E_HIP, "nk: parameter upload failed");
    return hipStreamSynchronize(st_) == hipSuccess ? MJX_OK : fail(MJX_E_HIP, "nk: sync failed");
  }

  // finish = false leaves the residual's finish to GMRES's first launch
  void residual(bool finish) {
    hipLaunchKernelGGL(k_nk_residual, dim3(RB), dim3(RT), 0, st_, (const double*)u_, u_old_, fold_, b_, a_, part_);
    if (finish) hipLaunchKernelGGL(k_nk_residual_final, dim3(1), dim3(RT), 0, st_, (const double*)part_, ctl_);
  }

  int read_ctl(GmresCtl* hc) {
    if (hipMemcpyAsync(hc, ctl_, sizeof(GmresCtl), hipMemcpyDeviceToHost, st_) != hipSuccess ||
        hipStreamSynchronize(st_) != hipSuccess)
      return fail(MJX_E_HIP, "nk: control block read-back failed");
    return MJX_OK;
  }

  // One Arnoldi step j of the current cycle (every kernel exits at once when
  // the cycle has already finished).
  GivensArrays givens_arrays() const {
    return GivensArrays{part_, h1_, h2_, H_, cs_, sn_, g_, ga_};
  }

  template <typename T>
  void launch_jvp(int j, int mode, T* t) {
    double* Vj = V_ + (long long)j * n_;
    const GivensArrays gva = givens_arrays();
    if (mode == JVP_GIVENS)
      hipLaunchKernelGGL((k_gm_jvp<T, JVP_GIVENS>), dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_,
                         (const double*)w_, ctl_, Vj, t, a_, gva, j - 1, c_.restart);
    else if (mode == JVP_START)
      hipLaunchKernelGGL((k_gm_jvp<T, JVP_START>), dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_,
                         (const double*)w_, ctl_, Vj, t, a_, gva, j - 1, c_.restart);
    else
      hipLaunchKernelGGL((k_gm_jvp<T, JVP_PLAIN>), dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_,
                         (const double*)w_, ctl_, Vj, t, a_, gva, j - 1, c_.restart);
  }

  // Arnoldi step j.  mode: JVP_GIVENS, its JVP makes the Givens step of
  // step j - 1; JVP_START, it makes the first cycle's start; givens_now:
  // step j's own Givens step is launched here (otherwise the next step's
  // JVP makes it).
  int arnoldi_step(int j, int mode, bool givens_now) {
    const int k = j + 1;
    if (c_.mv_dtype == MJX_F64) {
      launch_jvp<double>(j, mode, t64_);
      int rc = precond_t<double>(t64_, w_, t64b_, r2r64_, lamx64_, lamy64_);
      if (rc) return rc;
      hipLaunchKernelGGL(k_gm_dot1<double>, dim3(RB), dim3(RT), 0, st_, (const double*)V_, (const double*)w_, w_,
                         n_, k, part_, (const GmresCtl*)ctl_);
    } else {
      launch_jvp<float>(j, mode, t32_);
      int rc = precond32(t32_, w32_);
      if (rc) return rc;
      hipLaunchKernelGGL(k_gm_dot1<float>, dim3(RB), dim3(RT), 0, st_, (const double*)V_, (const float*)w32_, w_,
                         n_, k, part_, (const GmresCtl*)ctl_);
    }
    // partials: dot1 -> part_, upd_dot2 part_ -> part2_, upd_norm part2_ -> part_
    const GmresCtl* cc = ctl_;
    hipLaunchKernelGGL(k_gm_upd_dot2, dim3(RB), dim3(RT), 0, st_, (const double*)V_, w_, n_, k, (const double*)part_,
                       h1_, part2_, cc);
    hipLaunchKernelGGL(k_gm_upd_norm, dim3(RB), dim3(RT), 0, st_, (const double*)V_, w_, n_, k, (const double*)part2_,
                       h2_, part_, cc);
    if (givens_now) hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(RT), 0, st_, givens_arrays(), j, c_.restart, ctl_);
    return check_launch("arnoldi");
  }

  // x (FP64) -> J x (FP64), always in FP64 (used for restart residuals).
  int apply_A(const double* x, double* y) {
    hipLaunchKernelGGL(k_rd_jvp<double>, dim3(nblk(n_)), dim3(256), 0, st_, (const double*)u_, x, y, a_);
    return check_launch("jvp");
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
  GmresCtl* ctl_;
  double *H_, *cs_, *sn_, *g_, *h1_, *h2_, *ga_, *part_, *part2_;
  double *V_, *w_, *b_, *dx_, *xt_, *t64_, *t64b_, *r2r64_, *inst_;
  float *v32_, *t32_, *w32_, *t32b_, *r2r32_;
  const void *lamx64_, *lamy64_, *lamx32_, *lamy32_;
};

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
