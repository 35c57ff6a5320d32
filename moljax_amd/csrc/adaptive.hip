// Adaptive IMEX integration of a batch of Gray-Scott instances with a
// device-resident PI controller.
//
// One attempt of every instance (moljax/core/stepping.py:1146-1371 body):
//   y_full = S(dt_c) y, y_new = S(dt_c/2) S(dt_c/2) y    (step doubling,
//   stepping.py:566-608), err = y_new - y_full,
//   e = sqrt(mean over both fields of (err / (atol + rtol |y_new|))^2)
//   (scaled_error_norm, moljax/core/state.py:312-343), accept if e <= 1 and
//   y_new is finite, PI controller of moljax/core/dt_policy.py:178-268,544-610.
// is five steps on the device, with no host round trip:
//   chain(full) -> yf, chain(2 halves) -> yh    (gs_chain.h, shared kernels)
//   k_err_partial   per-instance FP64 partial sums and non-finite counts
//   k_controller    one wave per instance: fixed-order reduction, accept /
//                   reject, next dt, histories, instance tables for the next
//                   attempt
//   k_accept        y <- yh (interior) and the saved-state slot, for the
//                   instances that accepted
// Explicit schemes (forward Euler, SSPRK3, RK4 on the FD right-hand side,
// the explicit branch of moljax/core/stepping.py:723-1021) take the same
// attempt with explicit.hip's stage kernels in place of the IMEX chains.
// Each instance follows its own dt sequence; an instance that has left the
// loop (status != RUNNING) is skipped by the error, controller and accept
// kernels.  The host reads the per-instance state only between chunks of
// attempts, sized from the remaining (t_end - t) / dt.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "czt_api.h"
#include "explicit_api.h"
#include "gs_chain.h"
#include "mjx_common.h"
#include "moljax_hip.h"

namespace mjx {

namespace {

enum {
  ST_T = MJX_AS_T, ST_DT = MJX_AS_DT, ST_TEND = MJX_AS_TEND, ST_DTCFL = MJX_AS_DTCFL, ST_PREV = MJX_AS_PREV_ERR,
  ST_PREVPREV = MJX_AS_PREVPREV_ERR, ST_NACC = MJX_AS_N_ACCEPTED, ST_NREJ = MJX_AS_N_REJECTED,
  ST_WIDX = MJX_AS_WRITE_IDX, ST_STATUS = MJX_AS_STATUS, ST_ERR = MJX_AS_ERR, ST_ACCEPT = MJX_AS_ACCEPTED,
  ST_SLOT = MJX_AS_SLOT, ST_ATTEMPTS = MJX_AS_ATTEMPTS, ST_DTC = MJX_AS_DT_ATTEMPT
};
constexpr double RUNNING = 1, SUCCESS = 0, MAX_STEPS_REACHED = 2, DT_TOO_SMALL = 3, NON_FINITE_VALUES = 4;

// Attempt schemes: the IMEX chains of gs_chain.h, then the explicit methods.
enum { AD_IMEX_STRANG = 0, AD_IMEX_EULER = 1, AD_EXPL_EULER = 3, AD_EXPL_SSPRK3 = 4, AD_EXPL_RK4 = 5 };
__host__ __device__ inline bool is_explicit(int scheme) { return scheme >= AD_EXPL_EULER; }

struct Ctl {
  double atol, rtol, safety, kI, kP, dt_min, dt_max, max_factor, min_factor, order, dy, dx;
  long long save_every;
  int max_saves;
  int scheme;
};

constexpr int ERR_THREADS = 256;

// jnp.maximum / jnp.minimum propagate NaN (fmax/fmin do not).
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
__device__ __forceinline__ double nmin(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
__device__ __forceinline__ double nclip(double x, double lo, double hi) { return nmin(nmax(x, lo), hi); }

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Instance tables of the next attempt: dt_c = min(dt, t_end - t); the full
// step uses dt_c, the two halves dt_c / 2 (stepping.py:1216, :598-600).
__device__ void prepare_rows(const double* S, const double* model, InstParams* full, InstParams* half,
                             int scheme, double* dtc_out) {
  const double dtc = nmin(S[ST_DT], S[ST_TEND] - S[ST_T]);
  *dtc_out = dtc;
  const double Du = model[0], Dv = model[1], F = model[2], k = model[3];
  const double hf = dtc / 2.0;
  if (is_explicit(scheme)) {
    // explicit stages read D from the alpha columns
    full->beta_a = full->beta_b = half->beta_a = half->beta_b = 0.0;
    full->alpha_a = half->alpha_a = Du;
    full->alpha_b = half->alpha_b = Dv;
  } else {
    // Strang solves with dt/2 (stepping.py:399), Euler with dt (stepping.py:363).
    const double df = scheme == GS_STRANG ? hf : dtc;
    const double dh = scheme == GS_STRANG ? hf / 2.0 : hf;
    full->beta_a = df * Du;
    full->beta_b = df * Dv;
    half->beta_a = dh * Du;
    half->beta_b = dh * Dv;
    full->alpha_a = full->alpha_b = half->alpha_a = half->alpha_b = 0.0;
  }
  full->F = half->F = F;
  full->k = half->k = k;
  full->dt = dtc;
  half->dt = hf;
  full->aux = half->aux = 0.0;
}

__global__ void k_prepare(int B, double* state, const double* model, InstParams* full, InstParams* half,
                          int scheme, int max_saves) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* S = state + (long long)b * MJX_AS_STRIDE;
  S[ST_ACCEPT] = 0.0;
  S[ST_SLOT] = -1.0;
  // An instance whose loop condition already fails takes no attempt
  // (stepping.py:1208-1212) and gets its final status (:1356-1364).
  if (S[ST_STATUS] == RUNNING && !(S[ST_T] < S[ST_TEND] && S[ST_WIDX] < (double)max_saves)) {
    if (S[ST_T] >= S[ST_TEND]) S[ST_STATUS] = SUCCESS;
    else if (S[ST_WIDX] >= (double)max_saves) S[ST_STATUS] = MAX_STEPS_REACHED;
  }
  double dtc;
  prepare_rows(S, model + 4LL * b, full + b, half + b, scheme, &dtc);
  S[ST_DTC] = dtc;
}

// Per instance and block: sum of (e / (atol + rtol |y|))^2 over both fields
// and the number of non-finite y_new values, accumulated in FP64 in a fixed
// order (deterministic).
template <typename T>
__global__ __launch_bounds__(ERR_THREADS) void k_err_partial(const T* __restrict__ yf, const T* __restrict__ yh,
                                                             long long n, int nblk, double atol, double rtol,
                                                             const double* __restrict__ state,
                                                             double* __restrict__ partial) {
  const int b = blockIdx.x / nblk, blk = blockIdx.x - b * nblk;
  __shared__ double ls[ERR_THREADS / 64], lc[ERR_THREADS / 64];
  if (state[(long long)b * MJX_AS_STRIDE + ST_STATUS] != RUNNING) {
    if (threadIdx.x == 0) {
      partial[2LL * ((long long)b * nblk + blk)] = 0.0;
      partial[2LL * ((long long)b * nblk + blk) + 1] = 0.0;
    }
    return;
  }
  const T* f = yf + (long long)b * 2 * n;
  const T* h = yh + (long long)b * 2 * n;
  double acc = 0.0, bad = 0.0;
  for (long long i = (long long)blk * ERR_THREADS + threadIdx.x; i < 2 * n; i += (long long)nblk * ERR_THREADS) {
    const double yn = (double)h[i];
    const double e = yn - (double)f[i];
    const double q = e / (atol + rtol * fabs(yn));
    acc += q * q;
    bad += isfinite(yn) ? 0.0 : 1.0;
  }
  acc = wsum(acc);
  bad = wsum(bad);
  const int wv = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    ls[wv] = acc;
    lc[wv] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0, c = 0.0;
    for (int i = 0; i < ERR_THREADS / 64; ++i) {
      s += ls[i];
      c += lc[i];
    }
    partial[2LL * ((long long)b * nblk + blk)] = s;
    partial[2LL * ((long long)b * nblk + blk) + 1] = c;
  }
}

// One wave per instance.
__global__ __launch_bounds__(64) void k_controller(int nblk, double n_dof, const double* __restrict__ partial,
                                                   Ctl c, double* __restrict__ state,
                                                   const double* __restrict__ model, InstParams* full,
                                                   InstParams* half, double* t_hist, double* dt_hist) {
  const int b = blockIdx.x;
  double* S = state + (long long)b * MJX_AS_STRIDE;
  if (S[ST_STATUS] != RUNNING) {
    if (threadIdx.x == 0) {
      S[ST_ACCEPT] = 0.0;
      S[ST_SLOT] = -1.0;
    }
    return;
  }
  double s = 0.0, bad = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 64) {
    s += partial[2LL * ((long long)b * nblk + i)];
    bad += partial[2LL * ((long long)b * nblk + i) + 1];
  }
  s = wsum(s);
  bad = wsum(bad);
  if (threadIdx.x != 0) return;

  const double t = S[ST_T], t_end = S[ST_TEND], dtc = S[ST_DTC];
  const double e = nmax(sqrt(s / n_dof), 1e-10);
  const bool finite = bad == 0.0;
  const bool accept = (e <= 1.0) && finite;
  // propose_dt_imex (dt_policy.py:544-610)
  double dt_ad, new_prev = S[ST_PREV];
  if (e <= 1.0) {
    // pid_controller_dt (dt_policy.py:178-219)
    const double kI = c.kI / c.order, kP = c.kP / c.order;
    const double es = nmax(e, 1e-10), ps = nmax(S[ST_PREV], 1e-10);
    double factor = pow(1.0 / es, kI) * pow(ps / es, kP);
    factor = nclip(factor * c.safety, c.min_factor, c.max_factor);
    dt_ad = nclip(dtc * factor, c.dt_min, c.dt_max);
    new_prev = e;
  } else {
    // handle_rejected_step (dt_policy.py:222-268)
    const double kI = c.kI / c.order;
    const double es = nmax(e, 1.0);
    const double factor = nclip(c.safety * pow(1.0 / es, kI), c.min_factor, 1.0);
    dt_ad = nclip(dtc * factor, c.dt_min, c.dt_max);
  }
  const double dt_new = nclip(nmin(dt_ad, S[ST_DTCFL]), c.dt_min, c.dt_max);

  double status = S[ST_STATUS];
  double slot = -1.0;
  double t_next = t;
  double widx = S[ST_WIDX];
  if (accept) {
    const long long nacc = (long long)S[ST_NACC];
    if (nacc % c.save_every == 0) {
      const long long w = (long long)widx;
      t_hist[(long long)b * c.max_saves + w] = t + dtc;
      dt_hist[(long long)b * c.max_saves + w - 1] = dtc;
      slot = (double)w;
      widx += 1.0;
    }
    t_next = t + dtc;
    S[ST_DT] = dt_new;
    S[ST_PREVPREV] = S[ST_PREV];
    S[ST_PREV] = new_prev;
    S[ST_NACC] += 1.0;
  } else {
    const double dt_rej = nmax(dt_new * 0.5, c.dt_min);
    if (dt_rej <= c.dt_min * 1.1) status = DT_TOO_SMALL;
    S[ST_DT] = dt_rej;
    S[ST_NREJ] += 1.0;
  }
  if (!finite) status = NON_FINITE_VALUES;
  // Loop condition of the next attempt; final status as stepping.py:1356-1364.
  if (!(status == RUNNING && t_next < t_end && widx < (double)c.max_saves)) {
    if (t_next >= t_end) status = SUCCESS;
    else if (widx >= (double)c.max_saves) status = MAX_STEPS_REACHED;
  }
  S[ST_T] = t_next;
  S[ST_WIDX] = widx;
  S[ST_STATUS] = status;
  S[ST_ERR] = e;
  S[ST_ACCEPT] = accept ? 1.0 : 0.0;
  S[ST_SLOT] = slot;
  S[ST_ATTEMPTS] += 1.0;
  double dtn;
  prepare_rows(S, model + 4LL * b, full + b, half + b, c.scheme, &dtn);
  S[ST_DTC] = dtn;
}

// y <- yh on the interior of accepted instances, and the history slot.
template <typename T>
__global__ __launch_bounds__(256) void k_accept(const T* __restrict__ yh, T* __restrict__ u, T* __restrict__ v,
                                                long long pitch, long long istride, int ny, int nx, int nblk,
                                                const double* __restrict__ state, T* __restrict__ y_hist,
                                                int max_saves) {
  const int b = blockIdx.x / nblk, blk = blockIdx.x - b * nblk;
  const double* S = state + (long long)b * MJX_AS_STRIDE;
  if (S[ST_ACCEPT] == 0.0) return;
  const long long slot = (long long)S[ST_SLOT];
  const long long n = (long long)ny * nx;
  const T* h = yh + (long long)b * 2 * n;
  T* hu = (y_hist && slot >= 0) ? y_hist + ((long long)b * 2 * max_saves + slot) * n : nullptr;
  T* hv = hu ? hu + (long long)max_saves * n : nullptr;
  for (long long i = (long long)blk * 256 + threadIdx.x; i < n; i += (long long)nblk * 256) {
    const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
    const T a = h[i], c = h[n + i];
    u[(long long)b * istride + (long long)y * pitch + x] = a;
    v[(long long)b * istride + (long long)y * pitch + x] = c;
    if (hu) {
      hu[i] = a;
      hv[i] = c;
    }
  }
}

struct Layout {
  long long W, yf, yh, full, half, partial, xs, ymid, total;
  int nblk;
};

Layout layout(int dtype, int scheme, int B, int ny, int nx) {
  const long long es = dtype == MJX_F64 ? 8 : 4;
  const long long n = (long long)ny * nx;
  auto al = [](long long x) { return (x + 255) / 256 * 256; };
  Layout L;
  L.nblk = (int)std::min<long long>(256, std::max<long long>(1, (2 * n + 4095) / 4096));
  L.W = 0;
  L.yf = L.W + al((long long)B * ny * spec_work_cols(nx, ny) * 2 * es);
  L.yh = L.yf + al((long long)B * 2 * n * es);
  L.full = L.yh + al((long long)B * 2 * n * es);
  L.half = L.full + al((long long)B * sizeof(InstParams));
  L.partial = L.half + al((long long)B * sizeof(InstParams));
  L.xs = L.partial + al((long long)B * L.nblk * 2 * 8);
  L.ymid = L.xs + (is_explicit(scheme) ? al(explicit_scratch_elems(B, ny, nx) * es) : 0);
  L.total = L.ymid + (is_explicit(scheme) ? al((long long)B * 2 * n * es) : 0);
  return L;
}

template <typename T>
int adaptive_t(const Ctl& c, int B, int ny, int nx, T* u, T* v, long long pitch, long long istride,
               int fill_ghosts, const T* lamx, const T* lamy, const T* lam2d, const double* model, double* state,
               double* t_hist, double* dt_hist, T* y_hist, long long max_attempts, char* work, hipStream_t st,
               long long* attempts_done) {
  const Layout L = layout(sizeof(T) == 8 ? MJX_F64 : MJX_F32, c.scheme, B, ny, nx);
  cpx<T>* W = (cpx<T>*)(work + L.W);
  T* yf = (T*)(work + L.yf);
  T* yh = (T*)(work + L.yh);
  InstParams* full = (InstParams*)(work + L.full);
  InstParams* half = (InstParams*)(work + L.half);
  double* partial = (double*)(work + L.partial);
  T* xs = (T*)(work + L.xs);
  T* ymid = (T*)(work + L.ymid);
  const long long n = (long long)ny * nx;
  const int acc_blk = (int)std::min<long long>(64, (n + 1023) / 1024);

  hipLaunchKernelGGL(k_prepare, dim3((B + 63) / 64), dim3(64), 0, st, B, state, model, full, half, c.scheme,
                     c.max_saves);
  int rc = check_launch("adaptive_prepare");
  if (rc) return rc;

  std::vector<double> host((size_t)B * MJX_AS_STRIDE);
  long long done = 0;
  long long chunk = 4;
  for (;;) {
    for (long long i = 0; i < chunk; ++i) {
      if (max_attempts > 0 && done >= max_attempts) break;
      if (is_explicit(c.scheme)) {
        const int m = c.scheme - AD_EXPL_EULER;
        const Layout2<T> ys{u, v, pitch, istride}, of{yf, yf + n, nx, 2 * n}, om{ymid, ymid + n, nx, 2 * n},
            oh{yh, yh + n, nx, 2 * n};
        rc = gs_explicit_step(m, B, ny, nx, ys, of, c.dy, c.dx, full, xs, st);
        if (!rc) rc = gs_explicit_step(m, B, ny, nx, ys, om, c.dy, c.dx, half, xs, st);
        if (!rc) rc = gs_explicit_step(m, B, ny, nx, om, oh, c.dy, c.dx, half, xs, st);
      } else {
        rc = gs_chain_t<T>(c.scheme, B, ny, nx, u, v, pitch, istride, yf, yf + n, nx, 2 * n, 0, lamx, lamy, lam2d,
                           full, 1, W, st);
        if (!rc)
          rc = gs_chain_t<T>(c.scheme, B, ny, nx, u, v, pitch, istride, yh, yh + n, nx, 2 * n, 0, lamx, lamy,
                             lam2d, half, 2, W, st);
      }
      if (rc) return rc;
      hipLaunchKernelGGL((k_err_partial<T>), dim3(B * L.nblk), dim3(ERR_THREADS), 0, st, yf, yh, n, L.nblk, c.atol,
                         c.rtol, state, partial);
      if ((rc = check_launch("adaptive_err"))) return rc;
      hipLaunchKernelGGL(k_controller, dim3(B), dim3(64), 0, st, L.nblk, (double)(2 * n), partial, c, state, model,
                         full, half, t_hist, dt_hist);
      if ((rc = check_launch("adaptive_controller"))) return rc;
      hipLaunchKernelGGL((k_accept<T>), dim3(B * acc_blk), dim3(256), 0, st, yh, u, v, pitch, istride, ny, nx,
                         acc_blk, state, y_hist, c.max_saves);
      if ((rc = check_launch("adaptive_accept"))) return rc;
      ++done;
    }
    if (hipMemcpyAsync(host.data(), state, host.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(MJX_E_HIP, "adaptive: reading the controller state failed");
    double remaining = 0.0;
    bool running = false;
    for (int b = 0; b < B; ++b) {
      const double* S = host.data() + (size_t)b * MJX_AS_STRIDE;
      if (S[ST_STATUS] != RUNNING) continue;
      running = true;
      const double r = (S[ST_TEND] - S[ST_T]) / S[ST_DT];
      remaining = std::max(remaining, std::isfinite(r) ? r : 1.0);
    }
    if (!running || (max_attempts > 0 && done >= max_attempts)) break;
    chunk = (long long)std::min(64.0, std::max(1.0, std::ceil(remaining)));
  }
  *attempts_done = done;
  if (fill_ghosts) {
    const int dt_ = sizeof(T) == 8 ? MJX_F64 : MJX_F32;
    rc = mjx_bc_fill2d(dt_, B, ny, nx, u, pitch, istride, 0, nullptr, 1.0, 1.0, st);
    if (rc) return rc;
    rc = mjx_bc_fill2d(dt_, B, ny, nx, v, pitch, istride, 0, nullptr, 1.0, 1.0, st);
    if (rc) return rc;
  }
  return MJX_OK;
}

}  // namespace

}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_gs_adaptive_workspace_bytes(int dtype, int scheme, int B, int ny, int nx) {
  if (B < 1 || ny < 1 || nx < 1) return 0;
  return layout(dtype, scheme, B, ny, nx).total;
}

int mjx_gs_imex_adaptive(int dtype, int scheme, int B, int ny, int nx, void* u, void* v, long long pitch,
                         long long istride, int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d,
                         const double* model, const double* ctl, double* state, double* t_hist, double* dt_hist,
                         void* y_hist, int max_saves, long long max_attempts, long long* attempts_done, void* work,
                         void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (scheme != AD_IMEX_STRANG && scheme != AD_IMEX_EULER && !(scheme >= AD_EXPL_EULER && scheme <= AD_EXPL_RK4))
    return fail(MJX_E_ARG, "adaptive scheme must be 0/1 (IMEX Strang/Euler) or 3/4/5 (explicit Euler/SSPRK3/RK4)");
  if (B > 65535) return fail(MJX_E_LIMIT, "at most 65535 instances per call");
  if (B < 1) return fail(MJX_E_ARG, "batch must be >= 1");
  if (!is_explicit(scheme) && spec_generic(nx, ny) && (!czt_len_ok(nx) || !czt_len_ok(ny)))
    return fail(MJX_E_SIZE, "grid " + std::to_string(ny) + "x" + std::to_string(nx) +
                                " not supported: nx and ny must be in [2, 2048], or both powers of two in "
                                "[16, 4096]");
  if (ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty grid");
  if (!u || !v || !model || !ctl || !state || !t_hist || !dt_hist || !work)
    return fail(MJX_E_ARG, "null pointer argument");
  if (!is_explicit(scheme) && !lam2d && (!lamx || !lamy)) return fail(MJX_E_ARG, "need lamx and lamy, or lam2d");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  if (max_saves < 1) return fail(MJX_E_ARG, "max_saves must be >= 1");
  Ctl c;
  c.atol = ctl[0];
  c.rtol = ctl[1];
  c.safety = ctl[2];
  c.kI = ctl[3];
  c.kP = ctl[4];
  c.dt_min = ctl[5];
  c.dt_max = ctl[6];
  c.max_factor = ctl[7];
  c.min_factor = ctl[8];
  c.order = ctl[9];
  c.save_every = (long long)ctl[10];
  c.dy = ctl[11];
  c.dx = ctl[12];
  if (is_explicit(scheme) && !(c.dx > 0 && c.dy > 0)) return fail(MJX_E_ARG, "explicit schemes need dx, dy > 0");
  c.max_saves = max_saves;
  c.scheme = scheme;
  if (c.save_every < 1) return fail(MJX_E_ARG, "save_every must be >= 1");
  if (!(c.order > 0)) return fail(MJX_E_ARG, "order must be > 0");
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  long long dummy = 0;
  if (!attempts_done) attempts_done = &dummy;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MJX_F32)
    return adaptive_t<float>(c, B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)lamx,
                             (const float*)lamy, (const float*)lam2d, model, state, t_hist, dt_hist, (float*)y_hist,
                             max_attempts, (char*)work, st, attempts_done);
  return adaptive_t<double>(c, B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)lamx,
                            (const double*)lamy, (const double*)lam2d, model, state, t_hist, dt_hist,
                            (double*)y_hist, max_attempts, (char*)work, st, attempts_done);
}

}  // extern "C"
