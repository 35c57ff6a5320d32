// Explicit Runge-Kutta steps of the periodic Gray-Scott model
//   F(y) = D lap_h y + R(y),  R = (-u v^2 + F(1-u), u v^2 - (F+k) v),
// with lap_h the 5-point Laplacian (moljax/core/operators.py:159-322; the
// periodic ghost ring of moljax/core/bc.py:202-215 becomes index wrap).
//
// One fused kernel per Runge-Kutta stage reads the stage point z at the five
// stencil points of both fields, the base state y at the centre, and writes
// the next stage point (and for RK4 the running k-sum), in the arithmetic
// order of moljax/core/stepping.py:60-167:
//   Euler   y + dt F(y)
//   SSPRK3  y1 = y + dt F(y); y2 = (3/4 y + 1/4 y1) + (dt/4) F(y1);
//           y' = (1/3 y + 2/3 y2) + (2/3 dt) F(y2)
//   RK4     z_{i+1} = y + c_i dt k_i, s = ((k1 + 2 k2) + 2 k3) + k4,
//           y' = y + (dt/6) s
// Instance rows carry alpha_a = Du, alpha_b = Dv, F, k and dt.
#include <hip/hip_runtime.h>

#include <string>

#include "explicit_api.h"
#include "mjx_common.h"
#include "moljax_hip.h"

namespace mjx {
namespace {

template <typename T>
struct Stage {
  const T* yu;
  const T* yv;
  long long yp, yi;  // base state
  const T* zu;
  const T* zv;
  long long zp, zi;  // stage point
  T* ou;
  T* ov;
  long long op, oi;  // stage output
  T* su;
  T* sv;             // RK4 k-sum, dense (pitch nx, instance stride 2 n, v at +n)
  const InstParams* ip;
  int ny, nx;
  double dy2, dx2;  // the reference divides by h^2 (operators.py d2_central)
};

template <typename T>
__device__ __forceinline__ T lap5(const T* z, long long p, int y, int x, int ny, int nx, T dy2, T dx2) {
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const int ym = y == 0 ? ny - 1 : y - 1, yq = y == ny - 1 ? 0 : y + 1;
  const T c = z[(long long)y * p + x];
  // d2_central axis 0 (y) + axis 1 (x), each (f+ - 2 f + f-) / h^2
  const T d2y = (z[(long long)yq * p + x] - (T)2 * c + z[(long long)ym * p + x]) / dy2;
  const T d2x = (z[(long long)y * p + xp] - (T)2 * c + z[(long long)y * p + xm]) / dx2;
  return d2y + d2x;
}

template <typename T, int METHOD, int STAGE>
__global__ __launch_bounds__(256) void k_gs_stage(Stage<T> a) {
  const long long n = (long long)a.ny * a.nx;
  const int inst = blockIdx.y;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int y = (int)(i / a.nx), x = (int)(i - (long long)y * a.nx);
  const InstParams P = a.ip[inst];
  const T Du = (T)P.alpha_a, Dv = (T)P.alpha_b, F = (T)P.F, kk = (T)P.k, dt = (T)P.dt;
  const T dy2 = (T)a.dy2, dx2 = (T)a.dx2;
  const T* zu = a.zu + inst * a.zi;
  const T* zv = a.zv + inst * a.zi;
  const T cu = zu[(long long)y * a.zp + x], cv = zv[(long long)y * a.zp + x];
  const T uvv = cu * cv * cv;
  // rhs = L + N with L = D lap (operators.py:644-677 order)
  const T fu = Du * lap5(zu, a.zp, y, x, a.ny, a.nx, dy2, dx2) + (-uvv + F * ((T)1 - cu));
  const T fv = Dv * lap5(zv, a.zp, y, x, a.ny, a.nx, dy2, dx2) + (uvv - (F + kk) * cv);
  const long long yo = inst * a.yi + (long long)y * a.yp + x;
  const T yu = a.yu[yo], yv = a.yv[yo];
  T ru, rv;
  if constexpr (METHOD == EXPL_EULER) {
    ru = yu + dt * fu;
    rv = yv + dt * fv;
  } else if constexpr (METHOD == EXPL_SSPRK3) {
    if constexpr (STAGE == 0) {
      ru = yu + dt * fu;
      rv = yv + dt * fv;
    } else if constexpr (STAGE == 1) {
      ru = ((T)0.75 * yu + (T)0.25 * cu) + ((T)0.25 * dt) * fu;
      rv = ((T)0.75 * yv + (T)0.25 * cv) + ((T)0.25 * dt) * fv;
    } else {
      const T third = (T)(1.0 / 3.0), two3 = (T)(2.0 / 3.0);
      ru = (third * yu + two3 * cu) + (two3 * dt) * fu;
      rv = (third * yv + two3 * cv) + (two3 * dt) * fv;
    }
  } else {  // RK4
    const long long so = inst * 2 * n + i;
    if constexpr (STAGE == 0) {
      a.su[so] = fu;
      a.sv[so] = fv;
    } else if constexpr (STAGE < 3) {
      a.su[so] = a.su[so] + (T)2 * fu;
      a.sv[so] = a.sv[so] + (T)2 * fv;
    }
    if constexpr (STAGE < 2) {
      ru = yu + ((T)0.5 * dt) * fu;
      rv = yv + ((T)0.5 * dt) * fv;
    } else if constexpr (STAGE == 2) {
      ru = yu + dt * fu;
      rv = yv + dt * fv;
    } else {
      ru = yu + (dt / (T)6) * (a.su[so] + fu);
      rv = yv + (dt / (T)6) * (a.sv[so] + fv);
    }
  }
  const long long oo = inst * a.oi + (long long)y * a.op + x;
  a.ou[oo] = ru;
  a.ov[oo] = rv;
}

template <typename T, int METHOD, int STAGE>
int stage_launch(const Stage<T>& s, int B, hipStream_t st) {
  const long long n = (long long)s.ny * s.nx;
  hipLaunchKernelGGL((k_gs_stage<T, METHOD, STAGE>), dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, st, s);
  return check_launch("gs_explicit_stage");
}

template <typename T, int METHOD>
int run_stages(Stage<T> s, int B, int nstages, const Layout2<T>& out, T* zA, T* zB, hipStream_t st) {
  // stage k reads z (the base state for k = 0) and writes the next stage
  // point into zA/zB alternately; the last stage writes the output.
  const long long n = (long long)s.ny * s.nx;
  int rc = MJX_OK;
  for (int k = 0; k < nstages; ++k) {
    const bool last = k == nstages - 1;
    T* nu = last ? out.u : ((k & 1) ? zB : zA);
    T* nv = last ? out.v : ((k & 1) ? zB + n : zA + n);
    s.ou = nu;
    s.ov = nv;
    s.op = last ? out.pitch : s.nx;
    s.oi = last ? out.istride : 2 * n;
    switch (k) {
      case 0: rc = stage_launch<T, METHOD, 0>(s, B, st); break;
      case 1: rc = stage_launch<T, METHOD, 1>(s, B, st); break;
      case 2: rc = stage_launch<T, METHOD, 2>(s, B, st); break;
      default: rc = stage_launch<T, METHOD, 3>(s, B, st); break;
    }
    if (rc) return rc;
    s.zu = nu;
    s.zv = nv;
    s.zp = s.op;
    s.zi = s.oi;
  }
  return MJX_OK;
}

template <typename T>
int explicit_step_t(int method, int B, int ny, int nx, const Layout2<T>& y, const Layout2<T>& out, double dy,
                    double dx, const InstParams* ip, T* scratch, hipStream_t st) {
  const long long n = (long long)ny * nx;
  Stage<T> s{};
  s.yu = y.u;
  s.yv = y.v;
  s.yp = y.pitch;
  s.yi = y.istride;
  s.zu = y.u;
  s.zv = y.v;
  s.zp = y.pitch;
  s.zi = y.istride;
  T* zA = scratch;
  T* zB = scratch + (long long)B * 2 * n;
  T* sum = scratch + (long long)B * 4 * n;
  s.su = sum;
  s.sv = sum + n;
  s.ip = ip;
  s.ny = ny;
  s.nx = nx;
  s.dy2 = dy * dy;
  s.dx2 = dx * dx;
  // zA / zB hold stage points of all B instances with instance stride 2n:
  // offsets are relative to the instance base inside the kernel, so the
  // pointers passed are the buffer starts.
  switch (method) {
    case EXPL_EULER: return run_stages<T, EXPL_EULER>(s, B, 1, out, zA, zB, st);
    case EXPL_SSPRK3: return run_stages<T, EXPL_SSPRK3>(s, B, 3, out, zA, zB, st);
    case EXPL_RK4: return run_stages<T, EXPL_RK4>(s, B, 4, out, zA, zB, st);
    default: return fail(MJX_E_ARG, "unknown explicit method");
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy2(Layout2<T> src, Layout2<T> dst, int ny, int nx) {
  const long long n = (long long)ny * nx;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int inst = blockIdx.y;
  const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
  const long long so = inst * src.istride + (long long)y * src.pitch + x;
  const long long d = inst * dst.istride + (long long)y * dst.pitch + x;
  dst.u[d] = src.u[so];
  dst.v[d] = src.v[so];
}

// n_steps explicit steps in place on (u, v): the steps alternate between two
// dense buffers and the result is copied back once.
template <typename T>
int explicit_inplace_t(int method, int B, int ny, int nx, T* u, T* v, long long pitch, long long istride,
                       int ghosts, double dy, double dx, const InstParams* ip, int n_steps, T* work, hipStream_t st) {
  const long long n = (long long)ny * nx;
  T* scratch = work;
  T* P = work + explicit_scratch_elems(B, ny, nx);
  T* Q = P + (long long)B * 2 * n;
  const Layout2<T> state{u, v, pitch, istride};
  Layout2<T> cur = state;
  for (int s = 0; s < n_steps; ++s) {
    T* d = (s & 1) ? Q : P;
    const Layout2<T> out{d, d + n, nx, 2 * n};
    int rc = explicit_step_t<T>(method, B, ny, nx, cur, out, dy, dx, ip, scratch, st);
    if (rc) return rc;
    cur = out;
  }
  hipLaunchKernelGGL((k_copy2<T>), dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, st, cur, state, ny, nx);
  int rc = check_launch("gs_explicit_copy");
  if (rc || !ghosts) return rc;
  const int dt_ = sizeof(T) == 8 ? MJX_F64 : MJX_F32;
  rc = mjx_bc_fill2d(dt_, B, ny, nx, u, pitch, istride, 0, nullptr, 1.0, 1.0, st);
  if (!rc) rc = mjx_bc_fill2d(dt_, B, ny, nx, v, pitch, istride, 0, nullptr, 1.0, 1.0, st);
  return rc;
}

}  // namespace

int gs_explicit_step(int method, int B, int ny, int nx, const Layout2<float>& y, const Layout2<float>& out,
                     double dy, double dx, const InstParams* ip, float* scratch, hipStream_t st) {
  return explicit_step_t<float>(method, B, ny, nx, y, out, dy, dx, ip, scratch, st);
}
int gs_explicit_step(int method, int B, int ny, int nx, const Layout2<double>& y, const Layout2<double>& out,
                     double dy, double dx, const InstParams* ip, double* scratch, hipStream_t st) {
  return explicit_step_t<double>(method, B, ny, nx, y, out, dy, dx, ip, scratch, st);
}

}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_gs_explicit_workspace_bytes(int dtype, int B, int ny, int nx) {
  if (B < 1 || ny < 1 || nx < 1) return 0;
  const long long es = dtype == MJX_F64 ? 8 : 4;
  return (explicit_scratch_elems(B, ny, nx) + 4LL * B * ny * nx) * es;
}

int mjx_gs_explicit(int dtype, int method, int B, int ny, int nx, void* u, void* v, long long pitch,
                    long long istride, int fill_ghosts, double dy, double dx, const double* inst, int n_steps,
                    void* work, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (method < EXPL_EULER || method > EXPL_RK4) return fail(MJX_E_ARG, "method must be 0 (Euler), 1 (SSPRK3), 2 (RK4)");
  if (B < 1 || ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty grid");
  if (B > 65535) return fail(MJX_E_LIMIT, "at most 65535 instances per call");
  if (!u || !v || !inst || !work) return fail(MJX_E_ARG, "null pointer argument");
  if (pitch < nx) return fail(MJX_E_ARG, "pitch smaller than nx");
  if (fill_ghosts && pitch < nx + 2) return fail(MJX_E_ARG, "ghost fill needs pitch >= nx + 2");
  if (!(dx > 0 && dy > 0)) return fail(MJX_E_ARG, "dx and dy must be > 0");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (n_steps == 0) return MJX_OK;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (dtype == MJX_F32)
    return explicit_inplace_t<float>(method, B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, dy, dx,
                                     ip, n_steps, (float*)work, st);
  return explicit_inplace_t<double>(method, B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, dy, dx,
                                    ip, n_steps, (double*)work, st);
}

}  // extern "C"
