/*
 * moljax_hip.h -- C ABI of the MI355X (gfx950) kernels behind the moljax_amd
 * Python API.  Plain pointers, sizes and a hipStream_t passed as void*; no
 * torch types.  Every function returns 0 on success or a negative MJX_E_* code
 * and leaves a message for mjx_last_error() (thread-local).  No C++ exception
 * crosses this boundary.  All device pointers must be valid on the current HIP
 * device; all work is enqueued on `stream` (NULL = the legacy default stream)
 * and the calls are asynchronous unless stated otherwise.
 *
 * Array conventions (reference moljax/core/grid.py:115-208): a 2-D field is
 * row-major with axis 0 = y (rows) and axis 1 = x (columns).  Pointers passed
 * as `*_in` / `*_out` point at the first INTERIOR element; `pitch` is the row
 * stride and `istride` the stride between batch instances, both in elements.
 * A ghost-padded (ny+2, nx+2) array is therefore passed as base + pitch + 1
 * with pitch = nx + 2.  dtype: MJX_F32 or MJX_F64 for every real array of a
 * call; complex work buffers hold interleaved (re, im) pairs of that dtype.
 *
 * Per-instance parameter table `inst`: a device array of B rows of 8 float64:
 *   [beta_a, beta_b, alpha_a, alpha_b, F, k, dt, 0]
 * The spectral multiplier applied to field f at wavenumber (kx, ky) is
 *   ((1 + alpha_f*lam) / (1 - beta_f*lam))^nrep / (nx*ny),
 * with lam = lamx[kx] + lamy[ky] (or lam2d[ky*nx + kx]) and |1 - beta*lam|
 * clamped below at 1e-14 as in moljax/core/fft_solvers.py:280-283.
 * Two further multipliers are selected by negative nrep:
 *   nrep = -1:  alpha_f * lam / (nx*ny)        (apply D*lap, alpha = D)
 *   nrep = -2:  exp(beta_f * lam) / (nx*ny)    (apply exp(dt D lap), beta = dt D)
 *   nrep = -3:  the first form with the power taken per instance from the
 *               instance row's aux column (a non-negative integer), so one
 *               launch advances each instance by its own number of steps
 * (moljax/core/fft_operators.py:110-149 matvec / exp_matvec).
 */
#ifndef MOLJAX_HIP_H
#define MOLJAX_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MJX_F32 0
#define MJX_F64 1

#define MJX_OK 0
#define MJX_E_ARG (-1)     /* invalid argument (shape, dtype, pointer) */
#define MJX_E_SIZE (-2)    /* grid size not supported by the kernels */
#define MJX_E_HIP (-3)     /* HIP runtime error */
#define MJX_E_LIMIT (-4)   /* request exceeds a fixed capacity */

/* Library version (major*10000 + minor*100 + patch). */
int mjx_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* mjx_last_error(void);

/* Fill the FFT twiddle tables on the current device (FP64 sincospi, rounded
 * once to FP32) and wait for completion.  Called implicitly by every entry
 * point on its first use on a device; idempotent. */
int mjx_init(void);

/* Number of bytes of complex work buffer needed by the periodic spectral
 * entry points for B instances of ny x nx in the given dtype:
 * B*ny*nx complex values when nx and ny both have a Stockham plan (powers of
 * two in [16, 4096] and the mixed-radix lengths 96, 100, 192, 200, 250, 384,
 * 500, 768, 1000; packed x-spectra), else B*ny*2*(nx/2+1) (the x-half-spectra
 * of the two fields, chirp-z path). */
long long mjx_spectral_workspace_bytes(int dtype, int B, int ny, int nx);

/* Periodic 2-D spectral multiply of up to two real fields per instance:
 *   out_f = F^-1[ m_f(kx, ky) * F[in_f] ]   for f in {a, b},
 * m_f as described above.  One row pass (x-FFT of a + i*b, split into the two
 * Hermitian half-spectra), one column pass (y-FFT, multiply, inverse y-FFT)
 * and one row pass back.  b_in may be NULL (treated as zero; b_out ignored).
 * fill_ghosts != 0 also writes the periodic ghost ring of the outputs
 * (moljax/core/bc.py:202-215 order).  lam2d != NULL selects a full (ny, nx)
 * symbol array instead of the separable lamx/lamy tables.
 * Replaces: moljax/core/fft_solvers.py:261-294 (solve_helmholtz_2d),
 *           fft_solvers.py:368-445 (solve_diffusion_fft_field,
 *           apply_diffusion_inverse_fft) on periodic 2-D grids, and the
 *           FFT-CN step of benchmarks/benchmark_vs_scipy.py:74-91
 *           (alpha = beta = dt*D/2, lam = -(kx^2+ky^2)).
 * Sizes: nx, ny both Stockham lengths (powers of two in [16, 4096], and 96,
 * 100, 192, 200, 250, 384, 500, 768, 1000 with radix-5/6/10/12 stages:
 * Stockham line FFTs), or any nx, ny in [2, 2048] (chirp-z line transforms,
 * the same results to rounding).  The same rule holds for mjx_gs_imex_strang,
 * mjx_gs_imex_strang_var and mjx_gs_imex.  The first call for a side the
 * chirp-z kernels serve evaluates and uploads that side's plan tables
 * (blocking copies, kept for the process): make it before any stream
 * capture. */
int mjx_spectral_pair(int dtype, int B, int ny, int nx,
                      const void* a_in, const void* b_in, long long in_pitch, long long in_istride,
                      void* a_out, void* b_out, long long out_pitch, long long out_istride,
                      int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d,
                      const double* inst, int nrep, void* work, void* stream);

/* Gray-Scott IMEX-Strang, n_steps fixed steps, in place on (u, v):
 *   y* = (I - dt/2 D lap)^-1 y ; Heun on R(u,v) = (-uv^2 + F(1-u), uv^2 - (F+k)v)
 *   y** = y* + dt/2 (R(y*) + R(y* + dt R(y*))) ; y_new = (I - dt/2 D lap)^-1 y**
 * with the FD symbol of the periodic 5-point Laplacian in lamx/lamy (or the
 * full (ny, nx) array lam2d when non-NULL).  The
 * trailing half-solve of step s and the leading half-solve of step s+1 are
 * applied as one spectral multiply by the squared denominator (identical
 * operator; one FFT round trip per step instead of two).  inst rows carry
 * beta_a = dt/2*Du, beta_b = dt/2*Dv, F, k, dt per instance.
 * Replaces: moljax/core/stepping.py:370-435 (imex_strang_step) driven by
 *           stepping.py:1374-1436 (integrate_imex_fixed_dt) for the model of
 *           moljax/core/model.py:529-584 (create_gray_scott_periodic_fft).
 * work: mjx_spectral_workspace_bytes(dtype, B, ny, nx) bytes. */
int mjx_gs_imex_strang(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch,
                       long long istride, int fill_ghosts, const void* lamx, const void* lamy,
                       const void* lam2d, const double* inst, int n_steps, void* work, void* stream);

/* Spectral multiplier tables of the jitted ETD kernels, elementwise in FP64
 * over the (ny, nx) eigenvalues lam (device; ny = 1 for one line), z = dt
 * lam: mode 0 phi_1, phi_2, phi_3 of lam itself (out0..out2); 1 ETD1 {e^z,
 * dt phi1}; 2 ETD2 {e^z, dt phi1 + dt phi2, -dt phi2}; 3 ETDRK4 {e^z, dt b1,
 * dt b2 (= dt b3), dt b4} with b1 = phi1 - 3 phi2 + 4 phi3, b2 = 2 phi2 -
 * 4 phi3, b4 = -phi2 + 4 phi3; series below |z| = 1e-4.  even != 0 writes
 * each table's even part (m(k) + m(-k))/2 (indices reversed modulo ny, nx),
 * the multiplier the real part of ifft(m fft(x)) applies to real x.
 * Replaces: moljax/core/jit_kernels.py:28-53 (phi1-3) and the multiplier
 * arithmetic of :60-180 (etd1/etd2/etdrk4 kernels). */
int mjx_phi_tables(int ny, int nx, const double* lam, double dt, int mode, int even, double* out0, double* out1,
                   double* out2, double* out3, void* stream);

/* mjx_gs_imex_strang without any allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The spectral workspace
 * is a fixed, library-owned device array of MJX_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends); requests whose
 * mjx_spectral_workspace_bytes(dtype, B, ny, nx) exceed it return
 * MJX_E_LIMIT.  Stockham sides only (powers of two in [16, 4096] and the
 * mixed-radix lengths of mjx_spectral_pair; MJX_E_SIZE otherwise).  mjx_init must have run on the device (MJX_E_ARG otherwise).
 * Calls no hipMalloc/hipFree and creates no events.  Every call on a GPU
 * shares that GPU's scratch: calls that overlap on one GPU (different
 * streams or threads) race; calls on one stream are ordered.  Results are
 * bitwise those of mjx_gs_imex_strang. */
#define MJX_NOALLOC_SCRATCH_BYTES (64LL << 20)
int mjx_gs_imex_strang_noalloc(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch,
                               long long istride, int fill_ghosts, const void* lamx, const void* lamy,
                               const void* lam2d, const double* inst, int n_steps, void* stream);

/* Gray-Scott IMEX with a choice of scheme, n_steps fixed steps in place:
 * scheme 0 Strang (identical to mjx_gs_imex_strang), 1 IMEX-Euler
 *   y_new = (I - dt D lap)^-1 (y + dt R(y))      (stepping.py:317-367),
 * 2 IMEX-SSPRK2  y_new = y/2 + E(E(y))/2 with E the IMEX-Euler map
 *   (stepping.py:438-495), 3 the splitting of the Gray-Scott benchmark
 *   (benchmarks/benchmark_gray_scott.py:268-289): y += dt/2 R(y);
 *   y^ <- e^{dt D lam} y^; y += dt/2 R(y), 4 Strang in increment form: the
 *   operator of scheme 0 with the state kept physical in (u, v) and each
 *   diffusion solve applied as y + (H^k - 1)(y - (1, 0)), so that only the
 *   increment passes through the transforms (their FP32 rounding then
 *   scales with the increment; power-of-two sides only, MJX_E_SIZE
 *   otherwise).  inst beta columns: dt/2*D for Strang (0, 4), dt*D for
 *   Euler, SSPRK2 and the benchmark splitting.  Replaces
 * moljax/core/stepping.py:317-495 for the model of
 * moljax/core/model.py:529-584.  work as mjx_gs_imex_strang. */
int mjx_gs_imex(int dtype, int scheme, int B, int ny, int nx, void* u, void* v, long long pitch,
                long long istride, int fill_ghosts, const void* lamx, const void* lamy, const void* lam2d,
                const double* inst, int n_steps, void* work, void* stream);

/* Explicit steps of the periodic Gray-Scott model on the 5-point FD
 * right-hand side F = D lap_h y + R(y), n_steps in place on (u, v):
 * method 0 forward Euler, 1 SSPRK3 (Shu-Osher), 2 classical RK4, each stage
 * one fused kernel (stencil of both fields + reaction + stage update).
 * inst rows: alpha_a = Du, alpha_b = Dv, F, k, dt.  Any grid size.
 * Replaces: moljax/core/stepping.py:60-167 (euler_step, ssprk3_step,
 * rk4_step) with model.rhs of moljax/core/model.py:76-118 for the model of
 * model.py:213-292 (periodic), and integrate_fixed_dt (stepping.py:1024-1143).
 * work: mjx_gs_explicit_workspace_bytes(dtype, B, ny, nx) bytes. */
long long mjx_gs_explicit_workspace_bytes(int dtype, int B, int ny, int nx);
int mjx_gs_explicit(int dtype, int method, int B, int ny, int nx, void* u, void* v, long long pitch,
                    long long istride, int fill_ghosts, double dy, double dx, const double* inst, int n_steps,
                    void* work, void* stream);

/* Per-instance adaptive-integrator state: MJX_AS_STRIDE doubles per
 * instance, device memory.  The caller initialises T = t0,
 * DT = min(dt0, dt_cfl), TEND, DTCFL, PREV_ERR = PREVPREV_ERR = 1,
 * WRITE_IDX = 1, STATUS = 1 (running), the rest 0. */
enum {
  MJX_AS_T = 0, MJX_AS_DT = 1, MJX_AS_TEND = 2, MJX_AS_DTCFL = 3, MJX_AS_PREV_ERR = 4,
  MJX_AS_PREVPREV_ERR = 5, MJX_AS_N_ACCEPTED = 6, MJX_AS_N_REJECTED = 7, MJX_AS_WRITE_IDX = 8,
  MJX_AS_STATUS = 9, MJX_AS_ERR = 10, MJX_AS_ACCEPTED = 11, MJX_AS_SLOT = 12, MJX_AS_ATTEMPTS = 13,
  MJX_AS_DT_ATTEMPT = 14, MJX_AS_STRIDE = 16
};

long long mjx_gs_adaptive_workspace_bytes(int dtype, int scheme, int B, int ny, int nx);

/* Adaptive IMEX integration of B Gray-Scott instances on the periodic grid,
 * each instance with its own accept/reject and dt sequence, in place on
 * (u, v).  Every attempt takes one step of dt_c = min(dt, t_end - t) and two
 * of dt_c/2 (step doubling), e = max(RMS over both interior fields of
 * (y_new - y_full)/(atol + rtol |y_new|), 1e-10) in FP64, accepts when
 * e <= 1 and y_new is finite, and proposes the next dt by the PI controller
 * (reject: never grows, then halved), limited by DTCFL; status codes
 * 0 success, 1 running, 2 max steps, 3 dt too small, 4 non-finite.
 * scheme 0 IMEX-Strang, 1 IMEX-Euler (spectral chains, power-of-two grids),
 * 3 forward Euler, 4 SSPRK3, 5 RK4 (explicit on the 5-point FD right-hand
 * side, periodic, any grid size; the explicit branch of stepping.py:723-1021).
 * model: device, B x 4 doubles {Du, Dv, F, k}.
 * ctl: host, 13 doubles {atol, rtol, safety, kI, kP, dt_min, dt_max,
 *   max_factor, min_factor, order, save_every, dy, dx}; order is the error
 *   order (compute_error_order, dt_policy.py:411-423), DTCFL carries
 *   imex_cfl_dt for IMEX schemes and heisenberg_cfl_dt for explicit ones.
 * t_hist, dt_hist: device, B x max_saves doubles (t_hist[b][0] = t0 set by
 *   the caller); y_hist: NULL or device (B, 2, max_saves, ny, nx) of dtype
 *   (slot 0 = the initial interior set by the caller).  An accepted attempt
 *   with n_accepted % save_every == 0 stores t, dt_c and y_new at WRITE_IDX.
 * Runs until no instance is running, or max_attempts attempts (> 0) have
 * been made (the state then resumes with another call).  attempts_done may
 * be NULL.  fill_ghosts != 0 fills the periodic ghost ring at the end.
 * Replaces: moljax/core/stepping.py:1146-1371 (adaptive_integrate_imex) with
 *   :566-608 (estimate_error_imex_doubling), moljax/core/state.py:312-343
 *   (scaled_error_norm), moljax/core/dt_policy.py:178-268,544-610
 *   (pid_controller_dt, handle_rejected_step, propose_dt_imex).
 * work: mjx_gs_adaptive_workspace_bytes(dtype, scheme, B, ny, nx) bytes. */
int mjx_gs_imex_adaptive(int dtype, int scheme, int B, int ny, int nx, void* u, void* v, long long pitch,
                         long long istride, int fill_ghosts, const void* lamx, const void* lamy,
                         const void* lam2d, const double* model, const double* ctl, double* state,
                         double* t_hist, double* dt_hist, void* y_hist, int max_saves,
                         long long max_attempts, long long* attempts_done, void* work, void* stream);

/* Batched launch of Gray-Scott instances of uneven square sizes (no
 * reference counterpart; the nearest reference code batches fields of one
 * grid, moljax/core/jit_kernels.py:228-262,356-531).  A plan groups the
 * instances by size: group g has B[g] instances of n[g] x n[g] in (u[g],
 * v[g]) with row pitch pitch[g] and instance stride istride[g], parameter
 * rows inst[g] (B[g] x 8 doubles, F, k, dt and for Strang beta = dt/2 D),
 * tables tables[2g], tables[2g+1] = {ETDRK4 coefficients (from
 * mjx_etdrk4_coeffs), NULL} or {lamx, lamy} for Strang, and workspace
 * work[g] (mjx_etdrk4_workspace_bytes / mjx_spectral_workspace_bytes).
 * scheme 0 ETDRK4, 1 IMEX-Strang, 2 the Gray-Scott benchmark's splitting
 * (benchmarks/benchmark_gray_scott.py:268-289; tables {lamx, lamy}, inst
 * beta = dt*D).  Sizes: any n in [2, 2048] (chirp-z kernels), and the
 * Stockham sides (ETDRK4 32..1024, the others 16..4096).
 * mjx_batch_run advances every instance by n_steps: the steps run as
 * replays of one hipGraph of steps_per_graph steps whose branches are the
 * groups (one forked stream per group), the remainder as direct launches.
 * Each instance's result is bitwise that of the instance advanced alone.
 * A plan is used from one thread at a time; destroy it before freeing the
 * buffers it names. */
int mjx_batch_create(int dtype, int scheme, int n_groups, const int* n, const int* B, void* const* u,
                     void* const* v, const long long* pitch, const long long* istride,
                     const void* const* tables, const double* const* inst, void* const* work,
                     int steps_per_graph, void** plan);
/* As mjx_batch_create with one scheme per group, so that Gray-Scott groups
 * and single-field Helmholtz groups share one plan and one graph.  Schemes
 * 0-2 as above; 3, 4, 5: B[g] fields u[g] (v[g] unused, may be NULL) each
 * advanced by implicit diffusion steps u <- (I - beta lap_h)^-1 u, beta =
 * dt D from column 0 of its inst row, with 3 the periodic 5-point FD symbol
 * (FFT; moljax/core/fft_solvers.py:261-294), 4 homogeneous Dirichlet on a
 * vertex grid (DST-I) and 5 homogeneous Neumann on a cell grid (DCT-II)
 * (moljax/core/fft_nonperiodic.py:308-355; the batched copy
 * moljax/core/jit_kernels.py:510-531); tables {lamx, lamy} are the 1-D
 * symbols, work[g] holds mjx_spectral_workspace_bytes (3) or B n^2 values
 * (4, 5).  Sizes as for mjx_r2r_solve2d / mjx_spectral_pair. */
int mjx_batch_create_mixed(int dtype, int n_groups, const int* schemes, const int* n, const int* B,
                           void* const* u, void* const* v, const long long* pitch, const long long* istride,
                           const void* const* tables, const double* const* inst, void* const* work,
                           int steps_per_graph, void** plan);
int mjx_batch_run(void* plan, int n_steps, int fill_ghosts, void* stream);
/* mjx_batch_destroy waits for the plan's device (pending replays of its
 * graph) before it releases the graph, streams and events; the caller's
 * buffers stay the caller's. */
int mjx_batch_destroy(void* plan);

/* Autotuner (no reference counterpart): times the candidate tiles (lines per
 * workgroup) of the scheme's row and column kernels on scratch data of B
 * instances of n x n (reps timed repetitions each, HIP events), keeps the
 * fastest in the process-wide tile table, and returns them in chosen[0]
 * (row) and chosen[1] (column).  scheme 0 ETDRK4, 1 IMEX-Strang, 2 the
 * benchmark splitting (the Strang kernels' tiles).  Sides the Stockham
 * kernels do not take (chirp-z kernels, fixed tiles) return chosen = {0, 0}
 * without timing anything.
 * mjx_tune_get / mjx_tune_set read and write that table (kind 0 spectral
 * row, 1 spectral column, 2 ETDRK4 row, 3 ETDRK4 column; lines 0 = default),
 * so callers can persist tuned tiles. */
int mjx_autotune(int dtype, int scheme, int n, int B, int reps, int* chosen, void* stream);
int mjx_tune_get(int kind, int dtype, int n);
int mjx_tune_set(int kind, int dtype, int n, int lines);

/* Reductions over B strided (ny, nx) interiors, FP64 accumulation in a
 * fixed order (bitwise reproducible), one result per instance in out[b]
 * (device doubles):
 *   MJX_RED_SUMSQ      sum a^2                      (tree_norm2 before sqrt)
 *   MJX_RED_DOT        sum a b                      (tree_vdot)
 *   MJX_RED_MAXABS     max |a|, NaN propagates      (tree_norm_inf)
 *   MJX_RED_SCALED_SQ  sum (a / (atol + rtol |b|))^2 (scaled_error_norm
 *                      before mean and sqrt)
 *   MJX_RED_NONFINITE  count of non-finite a        (is_finite)
 *   MJX_RED_SUM        sum a                        (sum_interior, mean_interior;
 *                                                    utils.py:236-262)
 * accumulate != 0 combines with the existing out[b] (sum, or max for
 * MAXABS), so several fields reduce into one value with successive calls.
 * work: mjx_reduce_workspace_elems(B) doubles.
 * Replaces: moljax/core/state.py:115-189 (tree_norm2, tree_norm_inf,
 * tree_vdot), state.py:312-343 (scaled_error_norm), moljax/core/utils.py:186-197
 * (is_finite). */
enum {
  MJX_RED_SUMSQ = 0,
  MJX_RED_DOT = 1,
  MJX_RED_MAXABS = 2,
  MJX_RED_SCALED_SQ = 3,
  MJX_RED_NONFINITE = 4,
  MJX_RED_SUM = 5
};
long long mjx_reduce_workspace_elems(int B);
int mjx_reduce2d(int dtype, int op, int B, int ny, int nx, const void* a, long long a_pitch, long long a_istride,
                 const void* b, long long b_pitch, long long b_istride, double atol, double rtol, double* out,
                 int accumulate, double* work, void* stream);

/* Batch total of the per-instance results of mjx_reduce2d (B doubles on
 * the device): the sum (the maximum for MJX_RED_MAXABS) in a fixed order,
 * then its square root when take_sqrt != 0, into total[0].  With the
 * per-field calls of mjx_reduce2d it gives tree_norm2 / tree_vdot /
 * tree_norm_inf of a whole batch in one more launch.
 * Replaces: the final sum and sqrt of moljax/core/state.py:115-189. */
int mjx_reduce_finish(int op, int B, const double* per_inst, int take_sqrt, double* total, void* stream);

/* IMEX-Strang in place where instance b takes n_b = inst[8b + 7] steps
 * (0 <= n_b <= max_steps) in one batched chain: the column and row passes of
 * step s leave an instance with s >= n_b untouched, so every instance's
 * arithmetic is that of mjx_gs_imex_strang with n_steps = n_b (instances
 * with n_b = 0 hold unspecified values afterwards).  For continuous
 * batching of requests with different remaining step counts (no reference
 * counterpart; serves moljax/core/stepping.py:370-435 per request). */
int mjx_gs_imex_strang_var(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch,
                           long long istride, int fill_ghosts, const void* lamx, const void* lamy,
                           const void* lam2d, const double* inst, int max_steps, void* work, void* stream);

/* Non-periodic 2-D Helmholtz / Poisson solve on B instances of ny x nx
 * interior points:  kind 0 homogeneous Dirichlet (DST-I in x and y, vertex
 * grid), kind 1 homogeneous Neumann (orthonormal DCT-II), kind 2
 * homogeneous Dirichlet on a cell-centred grid (DST-II; ghost = -interior as
 * moljax/core/bc.py:237-258 fills it; symbol -4/h^2 sin^2(pi k/(2N)),
 * k = 1..N at [k-1]).  Sizes: any ny, nx in [2, 2048]; DST-I with ny+1,
 * nx+1 and the others with ny, nx powers of two in [16, 1024] run the
 * Stockham kernels, other sizes the chirp-z kernels (also for every other
 * mjx_r2r_* entry point).
 *   u = T^-1[ m(kx, ky) T[rhs] ],  mmode 0: m = (1 + alpha lam)/(1 - beta lam),
 *   mmode 1 (Poisson): m = 1/lam (0 where |lam| < 1e-14),  lam = lamx + lamy,
 *   mmode 2: the mmode-0 multiplier raised to the power inst[8b + 7] per
 *   instance (that many backward-Euler steps in one solve)
 * with lamx (nx) / lamy (ny) the 1-D symbols indexed like the transform
 * coefficients (Dirichlet k = 1..N at [k-1], Neumann k = 0..N-1 at [k]);
 * beta = inst[8b + 0], alpha = inst[8b + 2] per instance b.  work: B*ny*nx
 * reals.  Extends moljax/core/fft_nonperiodic.py:247-355
 * (solve_{poisson,helmholtz}_{dirichlet,neumann}, 1-D there) to 2-D. */
int mjx_r2r_solve2d(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                    long long in_istride, void* out, long long out_pitch, long long out_istride,
                    const void* lamx, const void* lamy, const double* inst, int mmode, void* work,
                    void* stream);

/* The same solve (kind 1 or 2, mmode 0 with alpha = 0, or mmode 1) with the
 * inhomogeneous boundary conditions of moljax/core/bc.py:218-307 on the
 * interior cells: side_values = {left, right, bottom, top} (host doubles)
 * are Dirichlet values g (kind 2, ghost = 2 g - interior) or Neumann fluxes
 * q (kind 1, ghost = interior -/+ 2 h q).  The ghosts' share of the 5-point
 * Laplacian (2 g / h^2, or -/+ 2 q / h) is folded into the right-hand side
 * (times beta, or -1 for Poisson), so the result u satisfies
 * u - beta lap_h u = rhs (or lap_h u = rhs) with lap_h evaluated on u padded
 * by bc.py's ghost fill.  dy, dx: grid spacings. */
int mjx_r2r_solve2d_bc(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                       long long in_istride, void* out, long long out_pitch, long long out_istride, const void* lamx,
                       const void* lamy, const double* inst, int mmode, const double* side_values, double dy,
                       double dx, void* work, void* stream);

/* The same solve along nlines independent 1-D lines of n points (line
 * strides in elements), one parameter row inst[0..7] for all lines.
 * Replaces moljax/core/fft_nonperiodic.py:247-355 for 1-D input. */
int mjx_r2r_solve1d(int dtype, int kind, int nlines, int n, const void* in, long long in_stride, void* out,
                    long long out_stride, const void* lam, const double* inst, int mmode, void* stream);

/* 1-D lines with bc.py boundary conditions at both ends, side_values =
 * {left, right} (host doubles), as mjx_r2r_solve2d_bc. */
int mjx_r2r_solve1d_bc(int dtype, int kind, int nlines, int n, const void* in, long long in_stride, void* out,
                       long long out_stride, const void* lam, const double* inst, int mmode,
                       const double* side_values, double dx, void* stream);

/* Exponential Euler for diffusion on non-periodic lines (nlines lines of n
 * points, u and N read with in_stride, out written with out_stride):
 *   out = T^-1[ e^{dt eig} T u + dt phi1(dt eig) T N ],
 * phi1(z) = (e^z - 1)/z, or 1 + z/2 + z^2/6 for |z| < 1e-4; eig: the
 * diffusion eigenvalues D lam indexed like the transform coefficients; kind
 * 0 DST-I (vertex Dirichlet), 1 orthonormal DCT-II (Neumann), 2 DST-II (cell
 * Dirichlet).  Replaces moljax/core/fft_nonperiodic.py:363-437
 * (etd1_dirichlet, etd1_neumann); out must not alias u or N. */
int mjx_r2r_etd1(int dtype, int kind, int nlines, int n, const void* u, const void* nonlin, long long in_stride,
                 void* out, long long out_stride, const void* eig, double dt, void* stream);

/* Raw 1-D transforms of nlines lines, times `scale`:  kind 0: DST-I
 * X_k = sum_n x_n sin(pi (k+1)(n+1)/(n_len+1)) (its own inverse up to
 * 2/(n_len+1));  kind 1: unnormalised DCT-II (inverse = 0) or DCT-III
 * (inverse != 0);  kind 2: unnormalised DST-II
 * X_k = sum_n x_n sin(pi (k+1)(2n+1)/(2N)) or DST-III
 * x_n = sum_k X_k sin(pi (k+1)(2n+1)/(2N)).  Replaces fft_nonperiodic.py:97-133 (dst_I_fast,
 * idst_I_fast). */
int mjx_r2r_transform1d(int dtype, int kind, int inverse, int nlines, int n, const void* in, long long in_stride,
                        void* out, long long out_stride, double scale, void* stream);

/* ETDRK4 coefficient tables for periodic Gray-Scott: six real tables
 * E, E2, Q, f1, f2, f3 of ny*(nx+2) elements each (dtype), in the separated
 * spectral layout (row ky, column s = kx for u, nx/2+1+kx for v), evaluated
 * in FP64 by the M = 32 contour mean of benchmarks/benchmark_gray_scott.py:
 * 336-356 with z = dt * D_f * (lamx[kx] + lamy[ky]); lamx, lamy are device
 * float64 tables.  mjx_etdrk4_coeff_elems gives the element count. */
long long mjx_etdrk4_coeff_elems(int ny, int nx);
int mjx_etdrk4_coeffs(int dtype, int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv,
                      double dt, void* coef, void* stream);

/* The same tables with a choice of the f2 weight: form 0 the Cox-Matthews
 * weight 2 f2 (as mjx_etdrk4_coeffs; the b2 of moljax/core/fft_integrators
 * .py:110-115), form 1 the benchmark's combination as written
 * (benchmarks/benchmark_gray_scott.py:414-415, f2 without the factor 2). */
int mjx_etdrk4_coeffs_form(int dtype, int form, int ny, int nx, const double* lamx, const double* lamy, double Du,
                           double Dv, double dt, void* coef, void* stream);

/* Batched coupled Gray-Scott ETDRK4 (Cox-Matthews), n_steps in place on
 * (u, v) of B ny x nx instances: square n x n with n in {32, ..., 1024} on
 * the Stockham kernels, any other sides in [2, 2048] on the chirp-z kernels
 * (the workspace is then six planes of ny * (nx+2) complex values per
 * instance; the coefficient tables are the same); per-instance F, k
 * from inst rows; coef from mjx_etdrk4_coeffs.  Four row passes (inverse
 * x-FFT, N(u,v), forward x-FFT) and four column passes (y-FFT of the
 * stage's operands, stage combination, inverse y-FFT) per step; on the
 * Stockham kernels every operand is a packed n x n plane (four per
 * instance, 4 n^2 + 2 n complex values with ETD2's N^ plane).  Replaces the coupled ETDRK4 of
 * benchmarks/benchmark_gray_scott.py:369-423 and serves
 * moljax/core/fft_integrators.py:252-341 (etdrk4_step) for Gray-Scott.
 * work: mjx_etdrk4_workspace_bytes(dtype, B, ny, nx). */
long long mjx_etdrk4_workspace_bytes(int dtype, int B, int ny, int nx);
int mjx_gs_etdrk4(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                  int fill_ghosts, const void* coef, const double* inst, int n_steps, void* work, void* stream);

/* The remaining two methods of the Gray-Scott benchmark, periodic with the
 * pseudo-spectral symbol lam = lamx[kx] + lamy[ky] (device tables of
 * dtype), n_steps in place on (u, v) of B instances (ny, nx powers of two in
 * [32, 1024]); work: mjx_gs_bench_workspace_bytes(dtype, B, ny, nx).
 *
 * mjx_gs_rk4_spectral: classical RK4 of y' = D lap y + R(y)
 *   (benchmarks/benchmark_gray_scott.py:89-127); inst rows carry
 *   alpha_a = Du, alpha_b = Dv, F, k, dt.
 * mjx_gs_cn_fixed_point: the benchmark's Crank-Nicolson step
 *   (benchmark_gray_scott.py:162-229): from y1 = y0 iterate
 *   r = y1 - y0 - dt/2 (D lap(y1 + y0) + R(y1) + R(y0)),
 *   y1 += (I - dt/2 D lap)^-1 (-r) while max|r_u| + max|r_v| > tol and
 *   fewer than max_iter iterations ran (at least one).  inst_lap rows:
 *   alpha = D (beta 0); inst_pc rows: beta = dt/2 D (alpha 0); both carry
 *   F, k, dt, and their aux columns are overwritten (per-instance iteration
 *   masks).  iters_total (device int[B], may be NULL) accumulates the
 *   iteration counts. */
long long mjx_gs_bench_workspace_bytes(int dtype, int B, int ny, int nx);
int mjx_gs_rk4_spectral(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                        const void* lamx, const void* lamy, const double* inst, int n_steps, void* work,
                        void* stream);
int mjx_gs_cn_fixed_point(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                          const void* lamx, const void* lamy, double* inst_lap, double* inst_pc, double tol,
                          int max_iter, int n_steps, int* iters_total, void* work, void* stream);

/* Exponential Euler / exponential Adams-Bashforth-2 tables for periodic
 * Gray-Scott: three real tables E = e^z, dt phi1(z), dt phi2(z) of ny*(nx+2)
 * elements each, z = dt * D_f * (lamx[kx] + lamy[ky]) in the separated
 * spectral layout of mjx_etdrk4_coeffs, FP64 with the series of
 * moljax/core/fft_integrators.py:37-64 (_phi1, _phi2) for |z| < 1e-4.
 * mjx_etd_phi_coeff_elems gives the element count. */
long long mjx_etd_phi_coeff_elems(int ny, int nx);
int mjx_etd_phi_coeffs(int dtype, int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv,
                       double dt, void* coef, void* stream);

/* Batched coupled Gray-Scott ETD1 (order 1) or ETD2 (order 2), n_steps in
 * place on (u, v) of B ny x nx instances (sizes as mjx_gs_etdrk4), one row
 * pass and one column pass per step; coef from mjx_etd_phi_coeffs; work as
 * for mjx_gs_etdrk4.
 *   ETD1: u^ <- E u^ + dt phi1 N^            (fft_integrators.py:119-174)
 *   ETD2: ... + dt phi2 (N^ - N^_{n-1})      (fft_integrators.py:177-249)
 * prev (ETD2 only): 0 = no N_{n-1} (the first step is ETD1, as the
 * reference's N_prev = None), 1 = the physical N_{n-1} of both fields in
 * (nu_prev, nv_prev) with the state's pitch / istride / origin, 2 = the N^
 * kept in work by the previous call on the same work and batch. */
int mjx_gs_etd_exp(int dtype, int order, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                   int fill_ghosts, const void* coef, const double* inst, int prev, const void* nu_prev,
                   const void* nv_prev, int n_steps, void* work, void* stream);

/* Implicit reaction-diffusion step by Newton-Krylov on an interior (ny, nx)
 * array: bc 0 homogeneous Dirichlet on a vertex grid (zero outside), bc 1
 * periodic; u_t = D lap_h u + f(u), f = reaction[0] u^3 + reaction[1] u^2 +
 * reaction[2] u + reaction[3];
 *   R(u) = c u - u_old - theta dt (D lap u + f(u)) - fold_term   (fold_term
 *   may be NULL; for Crank-Nicolson it is (1 - theta) dt F(u_old); for
 *   variable-step BDF2 c = alpha0, theta dt = beta and u_old holds
 *   -(alpha1 y_n + alpha2 y_{n-1}), newton_krylov.py:406-448),
 *   J v = c v - theta dt (D lap v + f'(u) v)   (hand-written JVP).
 * u (float64, in/out) holds the initial guess and receives the solution.
 * Newton as moljax/core/newton_krylov.py:171-359 (convergence tested before
 * the update, max_backtrack residual trials, fallback damped step);
 * left-preconditioned restarted GMRES(restart) with FP64 CGS2
 * orthogonalisation on deterministic wave-shuffle dot products and FP64
 * Givens rotations; one GMRES iteration = one Arnoldi step, convergence when
 * the Givens residual <= krylov_tol * ||M b||, max_krylov restart cycles.
 * nk_iparams = {operator dtype (MJX_F32/MJX_F64), precond (0 none, 1 exact
 *   (I - pdt D lap)^-1 by DST-I or FFT, 2 Jacobi on I - pdt D lap),
 *   jacobi sweeps, max_newton, max_krylov, restart, max_backtrack};
 * nk_params = {newton_tol, krylov_tol, damping, backtrack_factor,
 *   min_residual_decrease, jacobi_omega, c, pdt}: c the residual's leading
 *   coefficient (1 for BE/CN), pdt the preconditioner context's dt (the
 *   step dt, as moljax/core/newton_krylov.py:214 passes it);
 * lam_tables = {lamx f64, lamy f64, lamx f32, lamy f32} device symbol tables
 *   of the exact preconditioner (Dirichlet: DST-I eigenvalues, periodic: FD).
 * stats_out = {converged, newton_iters, gmres_iters, final ||R||}.  Synchronous
 * (the host reads Krylov scalars each Arnoldi step).
 * work: mjx_nk_workspace_bytes(ny, nx, restart) bytes. */
long long mjx_nk_workspace_bytes(int ny, int nx, int restart);
int mjx_rd_newton_krylov(int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta,
                         const double* reaction, double* u, const double* u_old, const double* fold_term,
                         const void* const* lam_tables, const double* nk_params, const int* nk_iparams,
                         double* stats_out, void* work, void* stream);

/* The residual R(u) (float64) and the JVP J(u) v (v, w in dtype, u float64)
 * of the problem above, as single kernels. */
int mjx_rd_residual(int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta,
                    double lead, const double* reaction, const double* u, const double* u_old,
                    const double* fold_term, double* r, void* stream);
int mjx_rd_jvp(int dtype, int ny, int nx, int bc, double dy, double dx, double D, double dt, double theta,
               double lead, const double* reaction, const double* u, const void* v, void* w, void* stream);

/* Deterministic FP64 dot products out[i] = <V_i, w> of k vectors of length n
 * (rows of V, stride n), wave64 shuffle + LDS reduction over a fixed block
 * partition; partial: mjx_dot_partial_elems(k) doubles of scratch.
 * Replaces the reductions of moljax/core/state.py:115-189. */
long long mjx_dot_partial_elems(int k);
int mjx_dot_f64(long long n, int k, const double* V, const double* w, double* out, double* partial, void* stream);

/* Opt-in timing of the fused IMEX chain's hot kernels (kind 0: the fused
 * Gray-Scott row pass, kind 1: the column pass).  While enabled every such
 * launch is bracketed by HIP events on its stream.  Enabling resets the
 * totals; mjx_kernel_timing waits for the recorded events and returns the
 * summed milliseconds and the launch count.  Measurement support with no
 * counterpart in the reference (its timing is wall clock around
 * block_until_ready, benchmarks/benchmark_utils.py:112-149). */
int mjx_kernel_timing_enable(int on);
int mjx_kernel_timing(int kind, double* total_ms, long long* count);
/* Per-launch durations (ms, in launch order) of kernel `kind` since timing
 * was enabled: copies up to cap values into out (may be NULL) and returns
 * the number recorded, or a negative error code. */
long long mjx_kernel_timing_samples(int kind, double* out, long long cap);

/* Finite-difference operator on ghost-padded arrays; writes the whole padded
 * output (interior = stencil, ghost ring = 0).  in/out point at the first
 * interior element; pitches must be >= nx + 2.
 *   op 0 Laplacian (d2/dy2 + d2/dx2)       moljax/core/operators.py:311-322
 *   op 1 second derivative along axis      operators.py:271-308
 *   op 2 central first derivative          operators.py:159-197
 *   op 3 upwind first derivative, velocity vel (array, same layout rules) or
 *        vel_scalar when vel == NULL       operators.py:200-268
 * axis: 0 = y (rows), 1 = x (columns). */
int mjx_fd_op2d(int dtype, int op, int B, int ny, int nx, const void* in, long long in_pitch,
                long long in_istride, void* out, long long out_pitch, long long out_istride, int axis,
                double dx, double dy, const void* vel, long long vel_pitch, long long vel_istride,
                double vel_scalar, void* stream);

/* out = a x + b y + c over n contiguous elements (y may be NULL); out may
 * alias x or y.  The element-wise StateDict algebra of
 * moljax/core/state.py:27-81 (tree_add, tree_sub, tree_scale, tree_axpy). */
int mjx_axpby(int dtype, long long n, double a, const void* x, double b, const void* y, double c, void* out,
              void* stream);

/* n_iter weighted-Jacobi sweeps for (I - beta lap) x = r on whole padded
 * arrays (moljax/core/preconditioners.py:121-198, DiffusionPreconditioner):
 * x0 = r, x <- omega (r + beta lap x) / diag + (1 - omega) x with
 * diag = 1 + beta (2/dx^2 + 2/dy^2) and lap the 5-point Laplacian on the
 * interior, 0 on the ghost ring.  r, x point at the first interior element
 * (pitches >= nx + 2); work: B (ny+2)(nx+2) elements (n_iter > 1); x must not
 * alias r.  n_iter = 0 copies r. */
int mjx_jacobi2d(int dtype, int B, int ny, int nx, const void* r, long long r_pitch, long long r_istride, void* x,
                 long long x_pitch, long long x_istride, void* work, double beta, double dy, double dx, double omega,
                 int n_iter, void* stream);

/* Ghost-ring fill of a padded field in place.  kind 0 periodic, 1 Dirichlet
 * (ghost = 2 g - interior), 2 Neumann (ghost = interior -/+ 2 h flux);
 * side_values = {left, right, bottom, top} (NULL = zeros).  Replaces
 * moljax/core/bc.py:202-307 (_apply_bc_2d_{periodic,dirichlet,neumann}),
 * corners as the reference's x-then-y order produces them. */
int mjx_bc_fill2d(int dtype, int B, int ny, int nx, void* f, long long pitch, long long istride, int kind,
                  const double* side_values, double dx, double dy, void* stream);

/* Gray-Scott reaction on n contiguous points: ru = -u v^2 + F(1-u),
 * rv = u v^2 - (F+k) v.  Replaces moljax/core/operators.py:644-667. */
int mjx_gs_reaction(int dtype, long long n, const void* u, const void* v, void* ru, void* rv, double F,
                    double k, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MOLJAX_HIP_H */
