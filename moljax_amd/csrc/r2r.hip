// C-ABI entry points of the non-periodic (DST-I / DCT-II / DST-II) solves.
#include <hip/hip_runtime.h>

#include <string>

#include "czt_api.h"
#include "mjx_common.h"
#include "r2r_kernels.h"

namespace mjx {

int r2r_run(int kind, int orient, int mode, int L, R2RArgs a, int B, hipStream_t st, float);
int r2r_run(int kind, int orient, int mode, int L, R2RArgs a, int B, hipStream_t st, double);

// Public kinds: 0 DST-I (vertex Dirichlet), 1 DCT-II (cell Neumann), 2 DST-II
// (cell Dirichlet, run as the alternating-sign DCT path).
enum { KIND_DST2 = 2 };

static int run(int dtype, int kind, int orient, int mode, int L, R2RArgs a, int B, hipStream_t st) {
  if (kind == KIND_DST2) {
    kind = R2R_DCT;
    a.alt = 1;
  }
  if (dtype == MJX_F32) return r2r_run(kind, orient, mode, L, a, B, st, 0.0f);
  return r2r_run(kind, orient, mode, L, a, B, st, 0.0);
}

static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Line length for a transform of n points: DST-I needs n+1, DCT / DST-II n, a
// power of two in [16, 1024].
static int line_len(int kind, int n) {
  const int m = kind == R2R_DST ? n + 1 : n;
  if (!pow2(m) || m < 16 || m > 1024) return -1;
  return 2 * m;
}

static int check_common(int dtype, int kind) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (kind != R2R_DST && kind != R2R_DCT && kind != KIND_DST2)
    return fail(MJX_E_ARG, "kind must be 0 (vertex Dirichlet / DST-I), 1 (Neumann / DCT-II) or 2 (cell Dirichlet / DST-II)");
  return MJX_OK;
}

// Lengths outside the power-of-two rule run the chirp-z kernels, n in [2, 2048].
static bool len_ok(int kind, int n) { return line_len(kind, n) > 0 || czt_len_ok(n); }

static std::string size_msg(int n) {
  return "length " + std::to_string(n) + " not supported: n must be in [2, 2048]";
}

// One batch of nlines contiguous lines (line strides in elements).
static CztR2RCall line_call(int dtype, int kind, int mode, int nlines, int n, const void* in, long long in_stride,
                            void* out, long long out_stride, double scale) {
  return CztR2RCall{dtype, kind, mode, 1, nlines, n, in, 1, in_stride, 0, out, 1, out_stride, 0, scale};
}

// Ghost-ring contributions of bc.py's fill to the 5-point Laplacian of the
// boundary cells: Dirichlet (ghost = 2 g - interior) adds 2 g / h^2;
// Neumann (ghost = interior -/+ 2 h q) adds -2 q / h (left, bottom) and
// +2 q / h (right, top).
static void set_lift(R2RArgs& a, int kind, const double* sv, int nsides, double dy, double dx) {
  a.lift_on = 1;
  for (int i = 0; i < 4; ++i) {
    const double v = i < nsides ? sv[i] : 0.0;
    const double h = i < 2 ? dx : dy;
    a.lift[i] = kind == KIND_DST2 ? 2.0 * v / (h * h) : (i % 2 == 0 ? -2.0 : 2.0) * v / h;
  }
}

static R2RArgs make(const void* in, long long in_es, long long in_ls, long long in_is, void* out, long long out_es,
                    long long out_ls, long long out_is, int nlines) {
  R2RArgs a{};
  a.in = in;
  a.in_es = in_es;
  a.in_ls = in_ls;
  a.in_is = in_is;
  a.out = out;
  a.out_es = out_es;
  a.out_ls = out_ls;
  a.out_is = out_is;
  a.nlines = nlines;
  a.scale = 1.0;
  return a;
}

// The chirp-z form of solve2d: rows forward into work, columns apply in
// place (the multiplier carries both dimensions' inverse normalisation), rows
// inverse into the output.
static int solve2d_czt(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                       long long in_istride, void* out, long long out_pitch, long long out_istride, const void* lamx,
                       const void* lamy, const double* inst, int mmode, const double* side_values, double dy,
                       double dx, void* work, hipStream_t st) {
  const long long plane = (long long)ny * nx;
  R2RArgs lifted{};
  if (side_values) set_lift(lifted, kind, side_values, 4, dy, dx);
  CztR2RCall r{dtype, kind, R2R_FWD, B, ny, nx, rhs, 1, in_pitch, in_istride, work, 1, nx, plane, 1.0};
  r.inst = inst;
  r.mmode = mmode;
  r.lift = side_values ? lifted.lift : nullptr;
  int rc = czt_r2r(r, st);
  if (rc) return rc;
  CztR2RCall c{dtype, kind, R2R_APPLY, B, nx, ny, work, nx, 1, plane, work, nx, 1, plane, 1.0};
  c.lam_line = lamy;
  c.lam_cross = lamx;
  c.n_cross = nx;
  c.inst = inst;
  c.mmode = mmode;
  if ((rc = czt_r2r(c, st))) return rc;
  CztR2RCall v{dtype, kind, R2R_INV, B, ny, nx, work, 1, nx, plane, out, 1, out_pitch, out_istride, 1.0};
  return czt_r2r(v, st);
}

static int solve2d(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                   long long in_istride, void* out, long long out_pitch, long long out_istride, const void* lamx,
                   const void* lamy, const double* inst, int mmode, const double* side_values, double dy, double dx,
                   void* work, void* stream) {
  const int Lx = line_len(kind, nx), Ly = line_len(kind, ny);
  int rc = MJX_OK;
  if (B < 1 || !rhs || !out || !lamx || !lamy || !inst || !work) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (in_pitch < nx || out_pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (mmode != R2R_HELM && mmode != R2R_POISSON && mmode != R2R_HELM_POW)
    return fail(MJX_E_ARG, "mmode must be 0 (Helmholtz), 1 (Poisson) or 2 (Helmholtz power)");
  if ((rc = ensure_tables(nullptr))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const long long plane = (long long)ny * nx;
  if (Lx < 0 || Ly < 0) return solve2d_czt(dtype, kind, B, ny, nx, rhs, in_pitch, in_istride, out, out_pitch,
                                           out_istride, lamx, lamy, inst, mmode, side_values, dy, dx, work, st);
  // rows: forward transform along x into the work array
  R2RArgs a = make(rhs, 1, in_pitch, in_istride, work, 1, nx, plane, ny);
  if (side_values) {
    a.inst = inst;
    a.mmode = mmode;
    set_lift(a, kind, side_values, 4, dy, dx);
  }
  rc = run(dtype, kind, R2R_ROW, R2R_FWD, Lx, a, B, st);
  if (rc) return rc;
  // columns: transform along y, multiply, inverse along y (in place)
  R2RArgs c = make(work, nx, 1, plane, work, nx, 1, plane, nx);
  c.lam_line = lamy;
  c.lam_cross = lamx;
  c.n_cross = kind != R2R_DST ? nx : 0;
  c.inst = inst;
  c.mmode = mmode;
  rc = run(dtype, kind, R2R_COL, R2R_APPLY, Ly, c, B, st);
  if (rc) return rc;
  // rows: inverse transform along x into the output
  R2RArgs r = make(work, 1, nx, plane, out, 1, out_pitch, out_istride, ny);
  r.scale = kind == R2R_DST ? 2.0 / (nx + 1) : 1.0;
  return run(dtype, kind, R2R_ROW, R2R_INV, Lx, r, B, st);
}

}  // namespace mjx

using namespace mjx;

extern "C" {

int mjx_r2r_solve2d(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                    long long in_istride, void* out, long long out_pitch, long long out_istride, const void* lamx,
                    const void* lamy, const double* inst, int mmode, void* work, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  if (!len_ok(kind, nx) || !len_ok(kind, ny))
    return fail(MJX_E_SIZE, std::string("grid ") + std::to_string(ny) + "x" + std::to_string(nx) +
                                " not supported: ny and nx must be in [2, 2048]");
  return solve2d(dtype, kind, B, ny, nx, rhs, in_pitch, in_istride, out, out_pitch, out_istride, lamx, lamy, inst,
                 mmode, nullptr, 0.0, 0.0, work, stream);
}

int mjx_r2r_solve2d_bc(int dtype, int kind, int B, int ny, int nx, const void* rhs, long long in_pitch,
                       long long in_istride, void* out, long long out_pitch, long long out_istride, const void* lamx,
                       const void* lamy, const double* inst, int mmode, const double* side_values, double dy,
                       double dx, void* work, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  if (kind == R2R_DST) return fail(MJX_E_ARG, "boundary values need kind 1 (Neumann) or 2 (cell Dirichlet)");
  if (mmode != R2R_HELM && mmode != R2R_POISSON)
    return fail(MJX_E_ARG, "boundary values need mmode 0 (Helmholtz, alpha = 0) or 1 (Poisson)");
  if (!side_values || !(dx > 0.0) || !(dy > 0.0)) return fail(MJX_E_ARG, "side_values and positive dx, dy required");
  if (!len_ok(kind, nx) || !len_ok(kind, ny))
    return fail(MJX_E_SIZE, std::string("grid ") + std::to_string(ny) + "x" + std::to_string(nx) +
                                " not supported: ny and nx must be in [2, 2048]");
  return solve2d(dtype, kind, B, ny, nx, rhs, in_pitch, in_istride, out, out_pitch, out_istride, lamx, lamy, inst,
                 mmode, side_values, dy, dx, work, stream);
}

int mjx_r2r_solve1d_bc(int dtype, int kind, int nlines, int n, const void* in, long long in_stride, void* out,
                       long long out_stride, const void* lam, const double* inst, int mmode,
                       const double* side_values, double dx, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  if (kind == R2R_DST) return fail(MJX_E_ARG, "boundary values need kind 1 (Neumann) or 2 (cell Dirichlet)");
  if (mmode != R2R_HELM && mmode != R2R_POISSON)
    return fail(MJX_E_ARG, "boundary values need mmode 0 (Helmholtz, alpha = 0) or 1 (Poisson)");
  if (!side_values || !(dx > 0.0)) return fail(MJX_E_ARG, "side_values and positive dx required");
  const int L = line_len(kind, n);
  if (!len_ok(kind, n)) return fail(MJX_E_SIZE, size_msg(n));
  if (nlines < 1 || !in || !out || !lam || !inst) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (in_stride < n || out_stride < n) return fail(MJX_E_ARG, "line stride smaller than n");
  if ((rc = ensure_tables(nullptr))) return rc;
  R2RArgs a = make(in, 1, in_stride, 0, out, 1, out_stride, 0, nlines);
  a.lam_line = lam;
  a.inst = inst;
  a.mmode = mmode;
  set_lift(a, kind, side_values, 2, dx, dx);
  if (L < 0) {
    CztR2RCall c = line_call(dtype, kind, R2R_APPLY, nlines, n, in, in_stride, out, out_stride, 1.0);
    c.lam_line = lam;
    c.inst = inst;
    c.mmode = mmode;
    c.lift = a.lift;
    return czt_r2r(c, (hipStream_t)stream);
  }
  return run(dtype, kind, R2R_ROW, R2R_APPLY, L, a, 1, (hipStream_t)stream);
}

int mjx_r2r_solve1d(int dtype, int kind, int nlines, int n, const void* in, long long in_stride, void* out,
                    long long out_stride, const void* lam, const double* inst, int mmode, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  const int L = line_len(kind, n);
  if (!len_ok(kind, n)) return fail(MJX_E_SIZE, size_msg(n));
  if (nlines < 1 || !in || !out || !lam || !inst) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (in_stride < n || out_stride < n) return fail(MJX_E_ARG, "line stride smaller than n");
  if (mmode != R2R_HELM && mmode != R2R_POISSON && mmode != R2R_HELM_POW)
    return fail(MJX_E_ARG, "mmode must be 0 (Helmholtz), 1 (Poisson) or 2 (Helmholtz power)");
  if ((rc = ensure_tables(nullptr))) return rc;
  if (L < 0) {
    CztR2RCall c = line_call(dtype, kind, R2R_APPLY, nlines, n, in, in_stride, out, out_stride, 1.0);
    c.lam_line = lam;
    c.inst = inst;
    c.mmode = mmode;
    return czt_r2r(c, (hipStream_t)stream);
  }
  R2RArgs a = make(in, 1, in_stride, 0, out, 1, out_stride, 0, nlines);
  a.lam_line = lam;
  a.inst = inst;
  a.mmode = mmode;
  return run(dtype, kind, R2R_ROW, R2R_APPLY, L, a, 1, (hipStream_t)stream);
}

int mjx_r2r_etd1(int dtype, int kind, int nlines, int n, const void* u, const void* nonlin, long long in_stride,
                 void* out, long long out_stride, const void* eig, double dt, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  const int L = line_len(kind, n);
  if (!len_ok(kind, n)) return fail(MJX_E_SIZE, size_msg(n));
  if (nlines < 1 || !u || !nonlin || !out || !eig) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (in_stride < n || out_stride < n) return fail(MJX_E_ARG, "line stride smaller than n");
  if ((rc = ensure_tables(nullptr))) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (L < 0) {
    CztR2RCall c = line_call(dtype, kind, R2R_APPLY, nlines, n, u, in_stride, out, out_stride, 1.0);
    c.lam_line = eig;
    c.mmode = R2R_EXP;
    c.beta0 = dt;
    if ((rc = czt_r2r(c, st))) return rc;
    c.in = nonlin;
    c.mmode = R2R_PHI1;
    c.accumulate = 1;
    return czt_r2r(c, st);
  }
  // out = T^-1[e^{dt eig} T u], then out += T^-1[dt phi1(dt eig) T N]
  R2RArgs a = make(u, 1, in_stride, 0, out, 1, out_stride, 0, nlines);
  a.lam_line = eig;
  a.mmode = R2R_EXP;
  a.beta0 = dt;
  rc = run(dtype, kind, R2R_ROW, R2R_APPLY, L, a, 1, st);
  if (rc) return rc;
  R2RArgs b = make(nonlin, 1, in_stride, 0, out, 1, out_stride, 0, nlines);
  b.lam_line = eig;
  b.mmode = R2R_PHI1;
  b.beta0 = dt;
  b.accumulate = 1;
  return run(dtype, kind, R2R_ROW, R2R_APPLY, L, b, 1, st);
}

int mjx_r2r_transform1d(int dtype, int kind, int inverse, int nlines, int n, const void* in, long long in_stride,
                        void* out, long long out_stride, double scale, void* stream) {
  int rc = check_common(dtype, kind);
  if (rc) return rc;
  const int L = line_len(kind, n);
  if (!len_ok(kind, n)) return fail(MJX_E_SIZE, size_msg(n));
  if (nlines < 1 || !in || !out) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (in_stride < n || out_stride < n) return fail(MJX_E_ARG, "line stride smaller than n");
  if ((rc = ensure_tables(nullptr))) return rc;
  const int mode = (kind != R2R_DST && inverse) ? R2R_INV : R2R_FWD;
  if (L < 0)
    return czt_r2r(line_call(dtype, kind, mode, nlines, n, in, in_stride, out, out_stride, scale),
                   (hipStream_t)stream);
  R2RArgs a = make(in, 1, in_stride, 0, out, 1, out_stride, 0, nlines);
  a.scale = scale;
  return run(dtype, kind, R2R_ROW, mode, L, a, 1, (hipStream_t)stream);
}

}  // extern "C"
