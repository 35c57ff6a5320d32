// Host-side check driver for the library's C++ host code under
// AddressSanitizer (built by build.py's build_host_check with the sanitizer
// on host code only).  With --cpu it drives every entry point's argument
// validation and the size/table helpers, which need no GPU; without it, it
// also runs small workloads through the entry points whose host code plans,
// allocates or caches (chirp-z tables, batch plans and their graphs, the
// Newton-Krylov driver, the autotuner, the allocation-free entry point) and
// checks that they succeed with finite results.  Correctness of the numbers
// is the test suite's business; this driver looks for host memory errors.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "moljax_hip.h"

static int g_fail = 0;

#define CHECK(cond, what)                                                        \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::printf("FAIL %s:%d %s (%s)\n", __FILE__, __LINE__, what, mjx_last_error()); \
      ++g_fail;                                                                  \
    }                                                                            \
  } while (0)

static void* P = reinterpret_cast<void*>(64);  // non-null stand-in, never dereferenced on error paths

static void cpu_checks() {
  CHECK(mjx_version() > 0, "version");
  CHECK(mjx_spectral_workspace_bytes(MJX_F32, 2, 64, 32) == 2LL * 64 * 32 * 8, "spectral workspace pow2");
  CHECK(mjx_spectral_workspace_bytes(MJX_F32, 2, 50, 30) == 2LL * 50 * 32 * 8, "spectral workspace any-length");
  CHECK(mjx_etdrk4_workspace_bytes(MJX_F32, 2, 64, 64) == 8LL * 2 * (4 * 64 * 64 + 2 * 64), "etd workspace");
  CHECK(mjx_gs_imex(MJX_F32, 4, 1, 96, 96, nullptr, nullptr, 98, 98 * 98, 1, nullptr, nullptr, nullptr, nullptr, 1,
                    nullptr, nullptr) == MJX_E_SIZE, "increment form takes power-of-two sides");
  CHECK(mjx_reduce_finish(MJX_RED_SUMSQ, 0, nullptr, 1, nullptr, nullptr) == MJX_E_ARG, "reduce finish empty");
  CHECK(mjx_nk_workspace_bytes(64, 64, 30) > 31LL * 64 * 64 * 8, "nk workspace");
  CHECK(mjx_reduce_workspace_elems(3) == 3 * 64, "reduce workspace");
  CHECK(mjx_etdrk4_coeff_elems(64, 64) == 6LL * 64 * 66, "coeff elems");
  CHECK(mjx_tune_set(0, 0, 128, 4) == MJX_OK && mjx_tune_get(0, 0, 128) == 4, "tune set/get");
  CHECK(mjx_tune_set(0, 0, 128, 3) == MJX_E_ARG, "tune rejects non-power-of-two");
  CHECK(mjx_tune_set(0, 0, 128, 0) == MJX_OK, "tune reset");
  // argument validation returns before any device work
  CHECK(mjx_spectral_pair(MJX_F32, 1, 3000, 32, P, nullptr, 32, 32 * 3000, P, nullptr, 32, 32 * 3000, 0, P, P,
                          nullptr, (const double*)P, 1, P, nullptr) == MJX_E_SIZE, "spectral pair size");
  CHECK(std::strstr(mjx_last_error(), "[2, 2048]") != nullptr, "size message");
  CHECK(mjx_gs_imex_strang(MJX_F32, 1, 2049, 2049, P, P, 2049, 2049LL * 2049, 0, P, P, nullptr, (const double*)P, 1,
                           P, nullptr) == MJX_E_SIZE, "strang size");
  CHECK(mjx_gs_imex_strang_noalloc(MJX_F32, 129, 256, 256, P, P, 258, 258LL * 258, 1, P, P, nullptr,
                                   (const double*)P, 3, nullptr) == MJX_E_LIMIT, "noalloc limit");
  CHECK(mjx_gs_imex_strang_noalloc(MJX_F32, 2, 252, 252, P, P, 254, 254LL * 254, 1, P, P, nullptr,
                                   (const double*)P, 3, nullptr) == MJX_E_SIZE, "noalloc size");
  CHECK(mjx_gs_imex(MJX_F32, 7, 1, 64, 64, P, P, 66, 66 * 66, 1, P, P, nullptr, (const double*)P, 1, P, nullptr) ==
            MJX_E_ARG, "imex scheme");
  void* plan = nullptr;
  const int n1[1] = {64}, b1[1] = {1};
  CHECK(mjx_batch_create(MJX_F32, 5, 1, n1, b1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 4,
                         &plan) == MJX_E_ARG && plan == nullptr, "batch scheme");
  const int s9[1] = {9};
  CHECK(mjx_batch_create_mixed(MJX_F32, 1, s9, n1, b1, &P, &P, nullptr, nullptr, nullptr, nullptr, nullptr, 4,
                               &plan) == MJX_E_ARG, "mixed batch scheme");
  const double sv[4] = {0.1, 0.2, 0.3, 0.4};
  CHECK(mjx_r2r_solve2d_bc(MJX_F64, 0, 1, 64, 64, P, 64, 64 * 64, P, 64, 64 * 64, P, P, (const double*)P, 0, sv,
                           0.1, 0.1, P, nullptr) == MJX_E_ARG, "r2r bc kind");
  CHECK(mjx_r2r_solve2d(MJX_F64, 2, 1, 2049, 64, P, 64, 64 * 2049, P, 64, 64 * 2049, P, P, (const double*)P, 0, P,
                        nullptr) == MJX_E_SIZE, "r2r size");
  CHECK(mjx_fd_op2d(MJX_F32, 3, 2, 64, 64, P, 66, 66 * 66, P, 66, 66 * 66, 1, 0.1, 0.1, P, 65, 66 * 66, 0.0,
                    nullptr) == MJX_E_ARG, "fd velocity pitch");
  CHECK(mjx_reduce2d(MJX_F32, 9, 1, 4, 4, P, 4, 16, nullptr, 4, 16, 0.0, 0.0, (double*)P, 0, (double*)P, nullptr) ==
            MJX_E_ARG, "reduce op");
  CHECK(mjx_phi_tables(1, 8, (const double*)P, 0.1, 7, 1, (double*)P, (double*)P, nullptr, nullptr, nullptr) ==
            MJX_E_ARG, "phi mode");
  CHECK(mjx_gs_etd_exp(MJX_F32, 3, 1, 64, 64, P, P, 64, 64 * 64, 0, P, (const double*)P, 0, P, P, 3, P, nullptr) ==
            MJX_E_ARG, "etd order");
  double st[4];
  const double nkp[8] = {1e-8, 1e-8, 1.0, 0.5, 1e-4, 0.6667, 1.0, 0.01};
  const int nki[7] = {MJX_F64, 1, 5, 10, 50, 300, 3};
  const double reac[4] = {-1, 0, 1, 0};
  const void* lt[4] = {P, P, P, P};
  CHECK(mjx_rd_newton_krylov(31, 31, 0, 0.03, 0.03, 1e-3, 1e-2, 1.0, reac, (double*)P, (const double*)P, nullptr, lt,
                             nkp, nki, st, P, nullptr) == MJX_E_LIMIT, "nk restart limit");
}

// ------------------------------------------------------------ GPU workloads
template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  explicit Dev(size_t count, const std::vector<T>* init = nullptr) : n(count) {
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) p = nullptr;
    if (p && init) (void)hipMemcpy(p, init->data(), count * sizeof(T), hipMemcpyHostToDevice);
    if (p && !init) (void)hipMemset(p, 0, count * sizeof(T));
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  std::vector<T> get() const {
    std::vector<T> h(n);
    (void)hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost);
    return h;
  }
};

template <typename T>
static bool finite(const std::vector<T>& v) {
  for (const T& x : v)
    if (!std::isfinite((double)x)) return false;
  return true;
}

static std::vector<double> fd_symbol(int n, double L) {
  std::vector<double> s(n);
  const double h = L / n;
  for (int i = 0; i < n; ++i) {
    const int f = i <= (n - 1) / 2 ? i : i - n;  // fftfreq ordering
    s[i] = (2.0 * std::cos(2.0 * M_PI * f / n) - 2.0) / (h * h);
  }
  return s;
}

static std::vector<double> dirichlet_symbol(int n) {
  std::vector<double> s(n);
  const double h = 1.0 / (n + 1);
  for (int k = 1; k <= n; ++k) s[k - 1] = -4.0 / (h * h) * std::pow(std::sin(M_PI * k / (2.0 * (n + 1))), 2);
  return s;
}

static std::vector<double> gs_field(int B, int W, int H, double base, double amp, unsigned seed) {
  std::vector<double> f((size_t)B * W * H);
  unsigned s = seed;
  for (auto& x : f) {
    s = s * 1664525u + 1013904223u;
    x = base + amp * ((s >> 8) / double(1 << 24));
  }
  return f;
}

template <typename T>
static std::vector<T> cast(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

template <typename T>
static void strang(int dtype, int n) {
  const int B = 2, W = n + 2;
  std::vector<double> row = {0.5 * 2e-5, 0.5 * 1e-5, 0, 0, 0.035, 0.065, 1.0, 0};
  std::vector<double> rows;
  for (int b = 0; b < B; ++b) rows.insert(rows.end(), row.begin(), row.end());
  auto hu = cast<T>(gs_field(B, W, W, 0.7, 0.3, 1)), hv = cast<T>(gs_field(B, W, W, 0.0, 0.25, 2));
  Dev<T> u(hu.size(), &hu), v(hv.size(), &hv);
  auto sym = cast<T>(fd_symbol(n, 2.5));
  Dev<T> lam(sym.size(), &sym);
  Dev<double> inst(rows.size(), &rows);
  Dev<char> work(mjx_spectral_workspace_bytes(dtype, B, n, n));
  const int rc = mjx_gs_imex_strang(dtype, B, n, n, u.p + W + 1, v.p + W + 1, W, (long long)W * W, 1, lam.p, lam.p,
                                    nullptr, inst.p, 3, work.p, nullptr);
  CHECK(rc == MJX_OK && finite(u.get()) && finite(v.get()), "IMEX-Strang");
  if (n == 64) {
    const int r2 = mjx_gs_imex_strang_noalloc(dtype, B, n, n, u.p + W + 1, v.p + W + 1, W, (long long)W * W, 1,
                                              lam.p, lam.p, nullptr, inst.p, 2, nullptr);
    CHECK(r2 == MJX_OK && finite(u.get()), "allocation-free IMEX-Strang");
  }
}

static void r2r(int kind, int n) {
  const int B = 2;
  auto rhs = gs_field(B, n, n, 0.0, 1.0, 3);
  Dev<double> in(rhs.size(), &rhs), out(rhs.size()), work(rhs.size());
  auto sym = kind == 0 ? dirichlet_symbol(n) : fd_symbol(n, 1.0);  // any real symbol serves the host paths
  Dev<double> lam(sym.size(), &sym);
  std::vector<double> rows = {0.01, 0, 0, 0, 0, 0, 0.1, 0, 0.02, 0, 0, 0, 0, 0, 0.1, 0};
  Dev<double> inst(rows.size(), &rows);
  const int rc = mjx_r2r_solve2d(MJX_F64, kind, B, n, n, in.p, n, (long long)n * n, out.p, n, (long long)n * n, lam.p,
                                 lam.p, inst.p, 0, work.p, nullptr);
  CHECK(rc == MJX_OK && finite(out.get()), "DST/DCT solve");
}

static void etdrk4(int n) {
  const int B = 2, dtype = MJX_F32;
  auto sym = fd_symbol(n, 1.0);
  Dev<double> lam(sym.size(), &sym);
  Dev<float> coef(mjx_etdrk4_coeff_elems(n, n));
  CHECK(mjx_etdrk4_coeffs(dtype, n, n, lam.p, lam.p, 2e-5, 1e-5, 1.0, coef.p, nullptr) == MJX_OK, "coefficients");
  auto hu = cast<float>(gs_field(B, n, n, 0.7, 0.3, 4)), hv = cast<float>(gs_field(B, n, n, 0.0, 0.25, 5));
  Dev<float> u(hu.size(), &hu), v(hv.size(), &hv);
  std::vector<double> rows = {0, 0, 0, 0, 0.04, 0.06, 1.0, 0, 0, 0, 0, 0, 0.03, 0.062, 1.0, 0};
  Dev<double> inst(rows.size(), &rows);
  Dev<char> work(mjx_etdrk4_workspace_bytes(dtype, B, n, n));
  const int rc = mjx_gs_etdrk4(dtype, B, n, n, u.p, v.p, n, (long long)n * n, 0, coef.p, inst.p, 2, work.p, nullptr);
  CHECK(rc == MJX_OK && finite(u.get()) && finite(v.get()), "ETDRK4");
}

static void mixed_batch() {
  // a Gray-Scott Strang group (64) and Helmholtz groups (periodic 40, Dirichlet 50)
  const int G = 3, dtype = MJX_F64;
  const int schemes[G] = {1, 3, 4}, ns[G] = {64, 40, 50}, Bs[G] = {2, 3, 2};
  std::vector<Dev<double>*> keep;
  void* u[G];
  void* v[G];
  long long pitch[G], istride[G];
  const void* tables[2 * G];
  const double* inst[G];
  void* work[G];
  for (int g = 0; g < G; ++g) {
    const int n = ns[g], B = Bs[g];
    auto f = gs_field(B, n, n, 0.6, 0.3, 10 + g);
    auto* du = new Dev<double>(f.size(), &f);
    auto* dv = new Dev<double>(f.size(), &f);
    auto sym = schemes[g] == 4 ? dirichlet_symbol(n) : fd_symbol(n, 1.0);
    auto* lam = new Dev<double>(sym.size(), &sym);
    std::vector<double> rows;
    for (int b = 0; b < B; ++b) {
      const double r[8] = {1e-4 * (b + 1), 5e-5, 0, 0, 0.035, 0.065, 1.0, 0};
      rows.insert(rows.end(), r, r + 8);
    }
    auto* ip = new Dev<double>(rows.size(), &rows);
    const long long wb = schemes[g] == 4 ? (long long)B * n * n * 8 : mjx_spectral_workspace_bytes(dtype, B, n, n);
    auto* wk = new Dev<double>((size_t)(wb / 8 + 1));
    keep.insert(keep.end(), {du, dv, lam, ip, wk});
    u[g] = du->p;
    v[g] = dv->p;
    pitch[g] = n;
    istride[g] = (long long)n * n;
    tables[2 * g] = lam->p;
    tables[2 * g + 1] = lam->p;
    inst[g] = ip->p;
    work[g] = wk->p;
  }
  void* plan = nullptr;
  int rc = mjx_batch_create_mixed(dtype, G, schemes, ns, Bs, u, v, pitch, istride, tables, inst, work, 2, &plan);
  CHECK(rc == MJX_OK && plan, "mixed batch create");
  if (plan) {
    CHECK(mjx_batch_run(plan, 5, 0, nullptr) == MJX_OK, "mixed batch run");
    CHECK(mjx_batch_run(plan, 3, 0, nullptr) == MJX_OK, "mixed batch rerun");
    CHECK(mjx_batch_destroy(plan) == MJX_OK, "mixed batch destroy");
  }
  (void)hipDeviceSynchronize();
  for (int g = 0; g < G; ++g) CHECK(finite(keep[5 * g]->get()), "mixed batch state");
  for (auto* d : keep) delete d;
}

static void newton_krylov(int n) {
  const double h = 1.0 / (n + 1);
  auto u0 = gs_field(1, n, n, 0.0, 0.5, 7);
  Dev<double> u(u0.size(), &u0), uold(u0.size(), &u0);
  auto sym = dirichlet_symbol(n);
  Dev<double> l64(sym.size(), &sym);
  auto sym32 = cast<float>(sym);
  Dev<float> l32(sym32.size(), &sym32);
  const void* lt[4] = {l64.p, l64.p, l32.p, l32.p};
  const double nkp[8] = {1e-8, 1e-8, 1.0, 0.5, 1e-4, 0.6667, 1.0, 1e-2};
  const double reac[4] = {-1, 0, 1, 0};
  for (int op : {MJX_F64, MJX_F32}) {
    const int nki[7] = {op, 1, 5, 10, 50, 30, 3};
    Dev<char> work(mjx_nk_workspace_bytes(n, n, 30));
    double st[4] = {0, 0, 0, 0};
    const int rc = mjx_rd_newton_krylov(n, n, 0, h, h, 1e-3, 1e-2, 1.0, reac, u.p, uold.p, nullptr, lt, nkp, nki, st,
                                        work.p, nullptr);
    CHECK(rc == MJX_OK && st[0] == 1.0 && finite(u.get()), "Newton-Krylov");
  }
}

static void misc(int n) {
  const int B = 2, W = n + 2;
  auto f = gs_field(B, W, W, 0.0, 1.0, 8);
  Dev<double> in(f.size(), &f), out(f.size());
  CHECK(mjx_fd_op2d(MJX_F64, 0, B, n, n, in.p + W + 1, W, (long long)W * W, out.p + W + 1, W, (long long)W * W, 0,
                    0.1, 0.1, nullptr, 0, 0, 0.0, nullptr) == MJX_OK && finite(out.get()), "FD Laplacian");
  const double sv[4] = {0.1, 0.2, 0.3, 0.4};
  CHECK(mjx_bc_fill2d(MJX_F64, B, n, n, in.p + W + 1, W, (long long)W * W, 1, sv, 0.1, 0.1, nullptr) == MJX_OK,
        "ghost fill");
  Dev<double> red(B), rw(mjx_reduce_workspace_elems(B));
  CHECK(mjx_reduce2d(MJX_F64, 0, B, n, n, in.p + W + 1, W, (long long)W * W, nullptr, 0, 0, 0, 0, red.p, 0, rw.p,
                     nullptr) == MJX_OK && finite(red.get()), "reduction");
  Dev<double> tot(1);
  CHECK(mjx_reduce_finish(MJX_RED_SUMSQ, B, red.p, 1, tot.p, nullptr) == MJX_OK && finite(tot.get()),
        "reduction batch total");
  auto sym = fd_symbol(n, 1.0);
  Dev<double> lam(sym.size(), &sym), o0(n), o1(n), o2(n), o3(n);
  CHECK(mjx_phi_tables(1, n, lam.p, 0.1, 3, 1, o0.p, o1.p, o2.p, o3.p, nullptr) == MJX_OK && finite(o3.get()),
        "phi tables");
  int chosen[2] = {0, 0};
  CHECK(mjx_autotune(MJX_F32, 1, 64, 4, 1, chosen, nullptr) == MJX_OK && chosen[0] > 0, "autotune");
  for (int k = 0; k < 4; ++k) (void)mjx_tune_set(k, 0, 64, 0);
}

int main(int argc, char** argv) {
  const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
  cpu_checks();
  if (!cpu_only) {
    CHECK(mjx_init() == MJX_OK, "init");
    strang<float>(MJX_F32, 64);
    strang<double>(MJX_F64, 48);  // any-length (chirp-z tables)
    r2r(0, 63);
    r2r(0, 50);
    r2r(1, 64);
    r2r(1, 37);
    etdrk4(64);
    etdrk4(50);
    mixed_batch();
    newton_krylov(31);
    newton_krylov(40);
    misc(64);
    CHECK(hipDeviceSynchronize() == hipSuccess, "device sync");
  }
  std::printf("host_check %s: %d failure(s)\n", cpu_only ? "cpu" : "gpu", g_fail);
  return g_fail == 0 ? 0 : 1;
}
