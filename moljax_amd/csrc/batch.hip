// Batched launch of many Gray-Scott instances of uneven grid sizes.
//
// A plan holds G groups; group g is B_g instances of one square size n_g
// (any side in [2, 2048]; power-of-two sides run the Stockham kernels),
// with its own state, parameter rows, coefficient/symbol tables and
// workspace (all caller-owned device memory).  run(n_steps):
//   1. pack every group into its spectral workspace (direct launches);
//   2. replay a hipGraph that advances every group by K steps: the capture
//      forks one stream per group off the plan's capture stream, so the
//      groups' kernels are independent branches of one graph and small
//      groups run alongside large ones instead of behind them;
//   3. the remainder (n_steps mod K) and the unpack run as direct launches.
// Schemes (per group): 0 ETDRK4 (tables: coefficients), 1 IMEX-Strang
// (tables: lamx, lamy), 2 the Gray-Scott benchmark's splitting (tables:
// lamx, lamy; inst beta = dt D; benchmarks/benchmark_gray_scott.py:268-289),
// and single-field Helmholtz groups whose step is one implicit diffusion
// solve u <- (I - beta lap_h)^-1 u with beta = dt D from the instance row
// (tables: lamx, lamy of the grid's 5-point Laplacian): 3 periodic (FFT),
// 4 homogeneous Dirichlet on a vertex grid (DST-I), 5 homogeneous Neumann
// on a cell grid (DCT-II) -- the solves of moljax/core/fft_solvers.py:261-294,
// moljax/core/fft_nonperiodic.py:308-355 and the batched copy of
// moljax/core/jit_kernels.py:510-531, repeated step by step.
// Kernel tiles (lines per workgroup) come from the autotuner table, which
// mjx_autotune fills by timing the candidate tiles on scratch data of the
// group's shape.  Results are bitwise those of the same instances advanced
// alone: every kernel is batch-invariant and the graph only changes the
// order of independent launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "czt_api.h"
#include "mjx_common.h"
#include "mjx_fft.h"
#include "spectral_api.h"

namespace mjx {

int etd_pack(int B, int n, const float* u, const float* v, long long pitch, long long istride, cpx<float>* work,
             hipStream_t st);
int etd_pack(int B, int n, const double* u, const double* v, long long pitch, long long istride,
             cpx<double>* work, hipStream_t st);
int etd_unpack(int B, int n, float* u, float* v, long long pitch, long long istride, int ghosts, cpx<float>* work,
               hipStream_t st);
int etd_unpack(int B, int n, double* u, double* v, long long pitch, long long istride, int ghosts,
               cpx<double>* work, hipStream_t st);
int etd_advance(int B, int n, const float* coef, const InstParams* ip, int n_steps, cpx<float>* work,
                hipStream_t st);
int etd_advance(int B, int n, const double* coef, const InstParams* ip, int n_steps, cpx<double>* work,
                hipStream_t st);


namespace {

enum {
  SCHEME_ETDRK4 = 0,
  SCHEME_STRANG = 1,
  SCHEME_BENCH_STRANG = 2,
  SCHEME_HELM_PERIODIC = 3,
  SCHEME_HELM_DIRICHLET = 4,
  SCHEME_HELM_NEUMANN = 5
};
bool helm_scheme(int s) { return s >= SCHEME_HELM_PERIODIC && s <= SCHEME_HELM_NEUMANN; }
// Strang-type chains run their first and last half-steps in pack / finish.
bool split_ends(int s) { return s == SCHEME_STRANG || s == SCHEME_BENCH_STRANG; }
int r2r_kind(int s) { return s == SCHEME_HELM_DIRICHLET ? 0 : 1; }

// Sizes of the Stockham ETDRK4 chain; other sides run the chirp-z one.
bool etd_stockham(int n) { return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024; }

struct Group {
  int scheme;
  int n, B;
  void *u, *v;
  long long pitch, istride;
  const void *t0, *t1;  // ETDRK4: coefficients; Strang: lamx, lamy
  const InstParams* ip;
  void* work;
};

template <typename T>
struct Ops {
  // One implicit diffusion solve per instance of a Helmholtz group, in place.
  static int helm_step(const Group& g, hipStream_t st) {
    constexpr int dt_ = sizeof(T) == 8 ? MJX_F64 : MJX_F32;
    const double* inst = reinterpret_cast<const double*>(g.ip);
    if (g.scheme == SCHEME_HELM_PERIODIC)
      return mjx_spectral_pair(dt_, g.B, g.n, g.n, g.u, nullptr, g.pitch, g.istride, g.u, nullptr, g.pitch,
                               g.istride, 0, g.t0, g.t1, nullptr, inst, 1, g.work, st);
    return mjx_r2r_solve2d(dt_, r2r_kind(g.scheme), g.B, g.n, g.n, g.u, g.pitch, g.istride, g.u, g.pitch, g.istride,
                           g.t0, g.t1, inst, 0, g.work, st);
  }
  static int pack(const Group& g, hipStream_t st) {
    if (helm_scheme(g.scheme)) {
      // the state stays in place; any-length sides get their chirp-z tables
      // now, before a graph capture replays the solves
      if (g.scheme == SCHEME_HELM_PERIODIC) return spec_generic(g.n, g.n) ? czt_prepare_periodic(g.n) : MJX_OK;
      return czt_prepare_r2r(r2r_kind(g.scheme), g.n);
    }
    if (g.scheme == SCHEME_ETDRK4) {
      if (!etd_stockham(g.n))
        return czt_etdrk4_pack(g.B, g.n, g.n, (const T*)g.u, (const T*)g.v, g.pitch, g.istride, (cpx<T>*)g.work, st);
      return etd_pack(g.B, g.n, (const T*)g.u, (const T*)g.v, g.pitch, g.istride, (cpx<T>*)g.work, st);
    }
    if (g.scheme == SCHEME_BENCH_STRANG) {
      // the first step up to its closing reaction half-step
      int rc = spec_row_fwd_pack(g.n, g.B, g.n, (const T*)g.u, (const T*)g.v, g.pitch, g.istride, (cpx<T>*)g.work,
                                 st, 1.0 / g.n);
      if (!rc) rc = spec_row_gs(g.n, g.B, g.n, (cpx<T>*)g.work, g.ip, st, RXN_HALF);
      if (!rc) rc = spec_col_apply(g.n, g.B, g.n, (cpx<T>*)g.work, (const T*)g.t0, (const T*)g.t1, nullptr, g.ip, -2, st);
      return rc;
    }
    int rc = spec_row_fwd_pack(g.n, g.B, g.n, (const T*)g.u, (const T*)g.v, g.pitch, g.istride, (cpx<T>*)g.work, st);
    if (rc) return rc;
    return spec_col_apply(g.n, g.B, g.n, (cpx<T>*)g.work, (const T*)g.t0, (const T*)g.t1, nullptr, g.ip, 1, st);
  }
  // Strang: `steps` middle steps (row_gs, col(H^2)); the last step's
  // trailing half-solve is applied by finish().
  // Benchmark splitting: each middle step closes the previous step's
  // reaction and opens its own (two half-steps in one row pass).
  static int advance(const Group& g, int steps, hipStream_t st) {
    if (helm_scheme(g.scheme)) {
      for (int s = 0; s < steps; ++s)
        if (int rc = helm_step(g, st)) return rc;
      return MJX_OK;
    }
    if (g.scheme == SCHEME_ETDRK4) {
      if (!etd_stockham(g.n))
        return czt_etdrk4_advance(g.B, g.n, g.n, (const T*)g.t0, g.ip, steps, (cpx<T>*)g.work, st);
      return etd_advance(g.B, g.n, (const T*)g.t0, g.ip, steps, (cpx<T>*)g.work, st);
    }
    const bool bench = g.scheme == SCHEME_BENCH_STRANG;
    for (int s = 0; s < steps; ++s) {
      int rc = spec_row_gs(g.n, g.B, g.n, (cpx<T>*)g.work, g.ip, st, bench ? RXN_HALF2 : RXN_HEUN);
      if (!rc)
        rc = spec_col_apply(g.n, g.B, g.n, (cpx<T>*)g.work, (const T*)g.t0, (const T*)g.t1, nullptr, g.ip,
                            bench ? -2 : 2, st);
      if (rc) return rc;
    }
    return MJX_OK;
  }
  static int finish(const Group& g, int ghosts, hipStream_t st) {
    if (helm_scheme(g.scheme)) return MJX_OK;
    if (g.scheme == SCHEME_ETDRK4) {
      if (!etd_stockham(g.n))
        return czt_etdrk4_unpack(g.B, g.n, g.n, (T*)g.u, (T*)g.v, g.pitch, g.istride, ghosts, (cpx<T>*)g.work, st);
      return etd_unpack(g.B, g.n, (T*)g.u, (T*)g.v, g.pitch, g.istride, ghosts, (cpx<T>*)g.work, st);
    }
    if (g.scheme == SCHEME_BENCH_STRANG)
      return spec_row_inv_unpack(g.n, g.B, g.n, (const cpx<T>*)g.work, (T*)g.u, (T*)g.v, g.pitch, g.istride, ghosts,
                                 st, 0.0, g.ip);
    int rc = spec_row_gs(g.n, g.B, g.n, (cpx<T>*)g.work, g.ip, st, RXN_HEUN);
    if (!rc) rc = spec_col_apply(g.n, g.B, g.n, (cpx<T>*)g.work, (const T*)g.t0, (const T*)g.t1, nullptr, g.ip, 1, st);
    if (!rc)
      rc = spec_row_inv_unpack(g.n, g.B, g.n, (const cpx<T>*)g.work, (T*)g.u, (T*)g.v, g.pitch, g.istride, ghosts, st);
    return rc;
  }
};

}  // namespace

struct BatchPlan {
  int dtype = MJX_F32, K = 1;
  int device = 0;         // the device the plan was created on
  bool launched = false;  // work of this plan was enqueued
  std::vector<Group> groups;
  hipStream_t cap = nullptr;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;  // ev[0] fork, ev[1 + g] join
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;

  // A plan may be destroyed right after an asynchronous run: wait for its
  // device first, so that no replay of the graph (nor a launch reading its
  // nodes) is still pending when the graph, its streams and events go.
  ~BatchPlan() {
    if (launched) {
      int cur = device;
      (void)hipGetDevice(&cur);
      if (cur != device) (void)hipSetDevice(device);
      (void)hipDeviceSynchronize();
      if (cur != device) (void)hipSetDevice(cur);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : ev) (void)hipEventDestroy(e);
    for (auto s : side) (void)hipStreamDestroy(s);
    if (cap) (void)hipStreamDestroy(cap);
  }


  // One graph advancing every group by K steps, one branch per group.
  int capture() {
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
      return fail(MJX_E_HIP, "batch: hipStreamCreate failed");
    const size_t G = groups.size();
    side.resize(G, nullptr);
    ev.resize(G + 1, nullptr);
    for (size_t g = 0; g < G; ++g)
      if (hipStreamCreateWithFlags(&side[g], hipStreamNonBlocking) != hipSuccess)
        return fail(MJX_E_HIP, "batch: hipStreamCreate failed");
    for (auto& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return fail(MJX_E_HIP, "batch: hipEventCreate failed");
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail(MJX_E_HIP, "batch: hipStreamBeginCapture failed");
    int rc = MJX_OK;
    if (hipEventRecord(ev[0], cap) != hipSuccess) rc = fail(MJX_E_HIP, "batch: event record in capture failed");
    for (size_t g = 0; g < G && rc == MJX_OK; ++g) {
      if (hipStreamWaitEvent(side[g], ev[0], 0) != hipSuccess) {
        rc = fail(MJX_E_HIP, "batch: stream wait in capture failed");
        break;
      }
      rc = dtype == MJX_F32 ? Ops<float>::advance(groups[g], K, side[g]) : Ops<double>::advance(groups[g], K, side[g]);
      if (rc == MJX_OK && (hipEventRecord(ev[1 + g], side[g]) != hipSuccess ||
                           hipStreamWaitEvent(cap, ev[1 + g], 0) != hipSuccess))
        rc = fail(MJX_E_HIP, "batch: join in capture failed");
    }
    hipGraph_t gph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cap, &gph);
    if (rc) {
      if (gph) (void)hipGraphDestroy(gph);
      return rc;
    }
    if (ec != hipSuccess || !gph) return fail(MJX_E_HIP, "batch: hipStreamEndCapture failed");
    graph = gph;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail(MJX_E_HIP, "batch: hipGraphInstantiate failed");
    return MJX_OK;
  }

  int run(int n_steps, int ghosts, hipStream_t st) {
    int rc = MJX_OK;
    launched = true;
    for (auto& g : groups) {
      rc = dtype == MJX_F32 ? Ops<float>::pack(g, st) : Ops<double>::pack(g, st);
      if (rc) return rc;
    }
    // Strang: the first step's leading half-solve is in pack(), the last
    // step's row pass and trailing half-solve in finish().  Benchmark
    // splitting: pack() runs the first step but its closing half-step,
    // which finish() applies.  Other schemes advance all n_steps in the
    // middle.  The graph advances every group K steps per replay; each
    // group takes its remaining middle steps as direct launches.
    int min_middle = n_steps;
    for (auto& g : groups) min_middle = std::min(min_middle, split_ends(g.scheme) ? n_steps - 1 : n_steps);
    const int reps = min_middle > 0 ? min_middle / K : 0;
    if (reps > 0 && !exec) {
      rc = capture();
      if (rc) return rc;
    }
    for (int r = 0; r < reps; ++r)
      if (hipGraphLaunch(exec, st) != hipSuccess) return fail(MJX_E_HIP, "batch: hipGraphLaunch failed");
    for (auto& g : groups) {
      const int rest = (split_ends(g.scheme) ? n_steps - 1 : n_steps) - reps * K;
      if (rest <= 0) continue;
      rc = dtype == MJX_F32 ? Ops<float>::advance(g, rest, st) : Ops<double>::advance(g, rest, st);
      if (rc) return rc;
    }
    for (auto& g : groups) {
      rc = dtype == MJX_F32 ? Ops<float>::finish(g, ghosts, st) : Ops<double>::finish(g, ghosts, st);
      if (rc) return rc;
    }
    return MJX_OK;
  }
};

// ------------------------------------------------------------ autotuner
namespace {

template <typename T>
int time_candidates(int kind, int n, int B, int reps, const std::vector<int>& cands, hipStream_t st,
                    const std::function<int()>& body, int* best_out, std::vector<float>* times) {
  constexpr int dt_ = sizeof(T) == 8 ? MJX_F64 : MJX_F32;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(MJX_E_HIP, "autotune: hipEventCreate failed");
  float best = 1e30f;
  int best_l = 0, rc = MJX_OK;
  for (int L : cands) {
    set_tuned_lines(kind, dt_, n, L);
    rc = body();  // warm-up (and first-launch code loading)
    if (rc) break;
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && rc == MJX_OK; ++r) rc = body();
    (void)hipEventRecord(e1, st);
    if (rc) break;
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(MJX_E_HIP, "autotune: hipEventSynchronize failed");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (times) times->push_back(ms / reps);
    if (ms < best) {
      best = ms;
      best_l = L;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_tuned_lines(kind, dt_, n, rc ? 0 : best_l);
  *best_out = rc ? 0 : best_l;
  (void)B;
  return rc;
}

// Default lines per workgroup (lines_default<T, N> at run time).
int lines_default_rt(int dtype, int n) {
  const int bt = dtype == MJX_F64 ? 128 : 256;
  const int tn = n / plan_p(n);
  return tn >= bt ? 1 : bt / tn;
}

std::vector<int> candidates(int dflt, int cap, int count) {
  std::vector<int> c;
  for (int L = std::min(dflt, cap); L >= 1 && (int)c.size() < count; L /= 2) c.push_back(L);
  return c;
}

template <typename T>
int autotune_t(int scheme, int n, int B, int reps, hipStream_t st, int* chosen) {
  const long long pts = (long long)B * n * n;
  T* u = nullptr;
  InstParams* ip = nullptr;
  void* work = nullptr;
  T* tab = nullptr;
  const long long work_bytes = mjx_etdrk4_workspace_bytes(sizeof(T) == 8 ? MJX_F64 : MJX_F32, B, n, n);
  const long long tab_elems = 6LL * n * (n + 2);
  int rc = MJX_OK;
  if (hipMalloc(&u, 2 * pts * sizeof(T)) != hipSuccess || hipMalloc(&ip, B * sizeof(InstParams)) != hipSuccess ||
      hipMalloc(&work, work_bytes) != hipSuccess || hipMalloc(&tab, tab_elems * sizeof(T)) != hipSuccess)
    rc = fail(MJX_E_HIP, "autotune: scratch allocation failed");
  if (!rc && (hipMemsetAsync(u, 0, 2 * pts * sizeof(T), st) != hipSuccess ||
              hipMemsetAsync(ip, 0, B * sizeof(InstParams), st) != hipSuccess ||
              hipMemsetAsync(work, 0, work_bytes, st) != hipSuccess ||
              hipMemsetAsync(tab, 0, tab_elems * sizeof(T), st) != hipSuccess))
    rc = fail(MJX_E_HIP, "autotune: scratch clear failed");
  const int dflt = lines_default_rt(sizeof(T) == 8 ? MJX_F64 : MJX_F32, n);
  if (!rc) {
    if (scheme == SCHEME_ETDRK4) {
      Group g{SCHEME_ETDRK4, n, B, u, u + pts, n, (long long)n * n, tab, nullptr, ip, work};
      auto body = [&]() { return Ops<T>::advance(g, 1, st); };
      rc = time_candidates<T>(TK_ETD_ROW, n, B, reps, candidates(dflt, n, 3), st, body, &chosen[0], nullptr);
      const int cm = etd_col_max(sizeof(T) == 8 ? MJX_F64 : MJX_F32, n);
      if (!rc) rc = time_candidates<T>(TK_ETD_COL, n, B, reps, candidates(cm, cm, 4), st, body, &chosen[1], nullptr);
    } else {
      cpx<T>* W = (cpx<T>*)work;
      auto row = [&]() { return spec_row_gs(n, B, n, W, ip, st, RXN_HEUN); };
      auto col = [&]() { return spec_col_apply(n, B, n, W, tab, tab, nullptr, ip, 2, st); };
      rc = time_candidates<T>(TK_ROW, n, B, reps, candidates(dflt, n, 4), st, row, &chosen[0], nullptr);
      if (!rc) rc = time_candidates<T>(TK_COL, n, B, reps, candidates(dflt, n / 2, 4), st, col, &chosen[1], nullptr);
    }
  }
  (void)hipStreamSynchronize(st);
  if (u) (void)hipFree(u);
  if (ip) (void)hipFree(ip);
  if (work) (void)hipFree(work);
  if (tab) (void)hipFree(tab);
  return rc;
}

}  // namespace
}  // namespace mjx

using namespace mjx;

extern "C" {

int mjx_batch_create_mixed(int dtype, int n_groups, const int* schemes, const int* n, const int* B, void* const* u,
                           void* const* v, const long long* pitch, const long long* istride,
                           const void* const* tables, const double* const* inst, void* const* work,
                           int steps_per_graph, void** plan) {
  if (!plan) return fail(MJX_E_ARG, "null plan pointer");
  *plan = nullptr;
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (n_groups < 1 || !schemes || !n || !B || !u || !v || !pitch || !istride || !tables || !inst || !work)
    return fail(MJX_E_ARG, "null pointer argument or no groups");
  if (steps_per_graph < 1 || steps_per_graph > 256) return fail(MJX_E_ARG, "steps_per_graph must be in [1, 256]");
  for (int g = 0; g < n_groups; ++g)
    if (schemes[g] < SCHEME_ETDRK4 || schemes[g] > SCHEME_HELM_NEUMANN)
      return fail(MJX_E_ARG, "batch group " + std::to_string(g) +
                                 ": scheme must be 0 (ETDRK4), 1 (Strang), 2 (benchmark splitting), 3 (periodic "
                                 "Helmholtz), 4 (Dirichlet Helmholtz) or 5 (Neumann Helmholtz)");
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  std::unique_ptr<BatchPlan> p(new BatchPlan());
  p->dtype = dtype;
  if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
  p->K = steps_per_graph;
  for (int g = 0; g < n_groups; ++g) {
    const int ng = n[g], scheme = schemes[g];
    const std::string gs = "batch group " + std::to_string(g);
    // Stockham sides run the Stockham kernels, other sides in [2, 2048] the
    // chirp-z ones (tables are created by the pack, before any capture)
    const bool ok = (scheme == SCHEME_ETDRK4 ? etd_stockham(ng) : supported_len(ng)) || czt_len_ok(ng);
    if (!ok) return fail(MJX_E_SIZE, gs + ": unsupported size " + std::to_string(ng));
    const bool two_fields = !helm_scheme(scheme);
    if (B[g] < 1 || !u[g] || (two_fields && !v[g]) || !inst[g] || !work[g] || !tables[2 * g] ||
        (scheme != SCHEME_ETDRK4 && !tables[2 * g + 1]))
      return fail(MJX_E_ARG, gs + ": null pointer or empty");
    if (pitch[g] < ng) return fail(MJX_E_ARG, gs + ": pitch smaller than n");
    Group gr{scheme, ng, B[g], u[g], two_fields ? v[g] : nullptr, pitch[g], istride[g], tables[2 * g],
             scheme != SCHEME_ETDRK4 ? tables[2 * g + 1] : nullptr, reinterpret_cast<const InstParams*>(inst[g]),
             work[g]};
    p->groups.push_back(gr);
  }
  *plan = p.release();
  return MJX_OK;
}

int mjx_batch_create(int dtype, int scheme, int n_groups, const int* n, const int* B, void* const* u, void* const* v,
                     const long long* pitch, const long long* istride, const void* const* tables,
                     const double* const* inst, void* const* work, int steps_per_graph, void** plan) {
  if (plan) *plan = nullptr;
  if (scheme != SCHEME_ETDRK4 && scheme != SCHEME_STRANG && scheme != SCHEME_BENCH_STRANG)
    return fail(MJX_E_ARG, "scheme must be 0 (ETDRK4), 1 (Strang) or 2 (benchmark splitting)");
  if (n_groups < 1 || n_groups > 4096) return fail(MJX_E_ARG, "n_groups must be in [1, 4096]");
  std::vector<int> schemes((size_t)n_groups, scheme);
  return mjx_batch_create_mixed(dtype, n_groups, schemes.data(), n, B, u, v, pitch, istride, tables, inst, work,
                                steps_per_graph, plan);
}

int mjx_batch_run(void* plan, int n_steps, int fill_ghosts, void* stream) {
  if (!plan) return fail(MJX_E_ARG, "null plan");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (n_steps == 0) return MJX_OK;
  BatchPlan* p = static_cast<BatchPlan*>(plan);
  if (fill_ghosts)
    for (auto& g : p->groups)
      if (!helm_scheme(g.scheme) && g.pitch < g.n + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= n + 2");
  return p->run(n_steps, fill_ghosts, (hipStream_t)stream);
}

int mjx_batch_destroy(void* plan) {
  delete static_cast<BatchPlan*>(plan);
  return MJX_OK;
}

int mjx_autotune(int dtype, int scheme, int n, int B, int reps, int* chosen, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (scheme != SCHEME_ETDRK4 && scheme != SCHEME_STRANG && scheme != SCHEME_BENCH_STRANG)
    return fail(MJX_E_ARG, "scheme must be 0 (ETDRK4), 1 (Strang) or 2 (benchmark splitting)");
  if (B < 1 || reps < 1 || !chosen) return fail(MJX_E_ARG, "B, reps >= 1 and chosen required");
  const bool stockham = scheme == SCHEME_ETDRK4 ? etd_stockham(n) : is_pow2_len(n);
  if (!stockham) {
    // the chirp-z kernels of other sides have fixed tiles: nothing to tune
    if (!czt_len_ok(n)) return fail(MJX_E_SIZE, "unsupported size " + std::to_string(n));
    chosen[0] = chosen[1] = 0;
    return MJX_OK;
  }
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return dtype == MJX_F32 ? autotune_t<float>(scheme, n, B, reps, st, chosen)
                          : autotune_t<double>(scheme, n, B, reps, st, chosen);
}

}  // extern "C"
