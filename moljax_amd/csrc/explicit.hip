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

#include <algorithm>
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

// Rows of a wave: each lane owns VE columns x_j = xs + 64 j + lane (every
// load instruction reads 64 consecutive elements, coalesced for any pitch
// or origin), and the wave walks R consecutive rows keeping the rows above,
// at and below the current one in registers, so each z value is read from
// memory once per wave (plus two halo rows per R).  Horizontal neighbours
// come from the neighbouring lane by a cross-lane permute; the two strip
// ends and the periodic wrap at x = nx - 1 are read from memory.
constexpr int STAGE_R = 16;  // rows per wave
constexpr int STAGE_LA = 2;  // rows loaded ahead of the one computed
constexpr int STAGE_WAVES = 4;

__device__ __forceinline__ int wrap_row(int r, int ny) {
  int w = r % ny;
  return w < 0 ? w + ny : w;
}

template <typename T, int VE>
struct Row {
  T a[VE];
};

// Columns x_j = xs + 64 j + lane of row `off`; FULL: the strip lies inside
// the row, so no lane needs a bound check.
template <bool FULL, typename T, int VE>
__device__ __forceinline__ void load_row(Row<T, VE>& r, const T* base, long long off, int xs, int lane, int nx) {
  const T* p = base + off + xs + lane;
#pragma unroll
  for (int j = 0; j < VE; ++j) r.a[j] = (FULL || xs + 64 * j + lane < nx) ? p[64 * j] : (T)0;
}

// Left and right neighbours of every owned column of row `c` (periodic in x).
template <bool FULL, typename T, int VE>
__device__ __forceinline__ void neighbours(const Row<T, VE>& c, const T* zrow, int xs, int lane, int nx, T* lft,
                                           T* rgt) {
  T up[VE], dn[VE];
#pragma unroll
  for (int j = 0; j < VE; ++j) {
    up[j] = __shfl(c.a[j], (lane + 63) & 63);  // lane - 1 (lane 0 gets lane 63)
    dn[j] = __shfl(c.a[j], (lane + 1) & 63);   // lane + 1 (lane 63 gets lane 0)
  }
  if constexpr (FULL) {
    // the strip's two halo values: lane 0 reads the column left of the
    // strip, lane 63 the column right of it, every other lane its own column
    const int xe = xs + 64 * VE;
    const int hx = lane == 0 ? (xs == 0 ? nx - 1 : xs - 1) : lane == 63 ? (xe == nx ? 0 : xe) : xs + lane;
    const T halo = zrow[hx];
#pragma unroll
    for (int j = 0; j < VE; ++j) {
      // lane 0 of column block j takes lane 63 of block j - 1, lane 63 lane 0 of block j + 1
      lft[j] = lane > 0 ? up[j] : (j > 0 ? up[j > 0 ? j - 1 : 0] : halo);
      rgt[j] = lane < 63 ? dn[j] : (j < VE - 1 ? dn[j < VE - 1 ? j + 1 : 0] : halo);
    }
  } else {
#pragma unroll
    for (int j = 0; j < VE; ++j) {
      const int x = xs + 64 * j + lane;
      lft[j] = lane > 0 ? up[j] : (j > 0 ? up[j > 0 ? j - 1 : 0] : (T)0);
      rgt[j] = lane < 63 ? dn[j] : (j < VE - 1 ? dn[j < VE - 1 ? j + 1 : 0] : (T)0);
      if (x < nx) {
        if (lane == 0 && j == 0) lft[j] = zrow[x == 0 ? nx - 1 : x - 1];
        if ((lane == 63 && j == VE - 1) || x == nx - 1) rgt[j] = zrow[x + 1 == nx ? 0 : x + 1];
      }
    }
  }
}

template <typename T>
struct StageCoef {
  T Du, Dv, F, kk, dt, idy2, idx2;
};

// One output row y from the rows above (m), at (c) and below (q) it.
template <typename T, int METHOD, int STAGE, int VE, bool SAME, bool FULL>
__device__ __forceinline__ void stage_row(const Stage<T>& a, const StageCoef<T>& k, int inst, int y, int xs, int lane,
                                          const Row<T, VE>& um, const Row<T, VE>& uc, const Row<T, VE>& uq,
                                          const Row<T, VE>& vm, const Row<T, VE>& vc, const Row<T, VE>& vq) {
  const int nx = a.nx;
  T ul[VE], ur[VE], vl[VE], vr[VE];
  neighbours<FULL>(uc, a.zu + inst * a.zi + (long long)y * a.zp, xs, lane, nx, ul, ur);
  neighbours<FULL>(vc, a.zv + inst * a.zi + (long long)y * a.zp, xs, lane, nx, vl, vr);
  const long long yo = inst * a.yi + (long long)y * a.yp + xs + lane;
  const long long oo = inst * a.oi + (long long)y * a.op + xs + lane;
  const long long so = inst * 2 * (long long)a.ny * nx + (long long)y * nx + xs + lane;
#pragma unroll
  for (int j = 0; j < VE; ++j) {
    if (!FULL && xs + 64 * j + lane >= nx) continue;
    const T cu = uc.a[j], cv = vc.a[j];
    const T uvv = cu * cv * cv;
    // d2_central axis 0 (y) + axis 1 (x), each (f+ - 2 f + f-) / h^2 (as a
    // product with 1/h^2); rhs = L + N with L = D lap (operators.py:644-677 order)
    const T lu = (uq.a[j] - (T)2 * cu + um.a[j]) * k.idy2 + (ur[j] - (T)2 * cu + ul[j]) * k.idx2;
    const T lv = (vq.a[j] - (T)2 * cv + vm.a[j]) * k.idy2 + (vr[j] - (T)2 * cv + vl[j]) * k.idx2;
    const T fu = k.Du * lu + (-uvv + k.F * ((T)1 - cu));
    const T fv = k.Dv * lv + (uvv - (k.F + k.kk) * cv);
    const T yu = SAME ? cu : a.yu[yo + 64 * j], yv = SAME ? cv : a.yv[yo + 64 * j];
    const T dt = k.dt;
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
      const long long si = so + 64 * j;
      if constexpr (STAGE == 0) {
        a.su[si] = fu;
        a.sv[si] = fv;
      } else if constexpr (STAGE < 3) {
        a.su[si] = a.su[si] + (T)2 * fu;
        a.sv[si] = a.sv[si] + (T)2 * fv;
      }
      if constexpr (STAGE < 2) {
        ru = yu + ((T)0.5 * dt) * fu;
        rv = yv + ((T)0.5 * dt) * fv;
      } else if constexpr (STAGE == 2) {
        ru = yu + dt * fu;
        rv = yv + dt * fv;
      } else {
        ru = yu + (dt / (T)6) * (a.su[si] + fu);
        rv = yv + (dt / (T)6) * (a.sv[si] + fv);
      }
    }
    a.ou[oo + 64 * j] = ru;
    a.ov[oo + 64 * j] = rv;
  }
}

// The wave's rows y0 .. y0 + rows - 1.  Slot i % RING of a ring of row
// registers holds row y0 - 1 + i; row r needs slots r, r + 1, r + 2, and
// while it is computed the loads of the next STAGE_LA rows are in flight.
// The loop is unrolled by RING, so the ring rotates by renaming registers.
template <typename T, int METHOD, int STAGE, int VE, bool SAME, bool FULL>
__device__ __forceinline__ void stage_rows(const Stage<T>& a, const StageCoef<T>& k, int inst, int y0, int rows,
                                           int xs, int lane) {
  constexpr int RING = 3 + STAGE_LA;
  const int nx = a.nx, ny = a.ny;
  const T* zu = a.zu + inst * a.zi;
  const T* zv = a.zv + inst * a.zi;
  const long long zp = a.zp;
  Row<T, VE> u[RING], v[RING];
#pragma unroll
  for (int i = 0; i < 2 + STAGE_LA; ++i) {
    if (i <= rows + 1) {
      const long long off = wrap_row(y0 - 1 + i, ny) * zp;
      load_row<FULL>(u[i], zu, off, xs, lane, nx);
      load_row<FULL>(v[i], zv, off, xs, lane, nx);
    }
  }
  for (int r0 = 0; r0 < rows; r0 += RING) {
#pragma unroll
    for (int c = 0; c < RING; ++c) {
      const int r = r0 + c;
      if (r < rows) {
        const int i = r + 2 + STAGE_LA;  // next row to fetch
        if (i <= rows + 1) {
          const long long off = wrap_row(y0 - 1 + i, ny) * zp;
          load_row<FULL>(u[(c + 2 + STAGE_LA) % RING], zu, off, xs, lane, nx);
          load_row<FULL>(v[(c + 2 + STAGE_LA) % RING], zv, off, xs, lane, nx);
        }
        stage_row<T, METHOD, STAGE, VE, SAME, FULL>(a, k, inst, y0 + r, xs, lane, u[c % RING], u[(c + 1) % RING],
                                                    u[(c + 2) % RING], v[c % RING], v[(c + 1) % RING],
                                                    v[(c + 2) % RING]);
      }
    }
  }
}

template <typename T, int METHOD, int STAGE, int VE, bool SAME>
__global__ __launch_bounds__(256) void k_gs_stage(Stage<T> a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xs = blockIdx.x * 64 * VE;
  const int y0 = (blockIdx.y * STAGE_WAVES + wave) * STAGE_R;
  if (y0 >= a.ny) return;
  const int rows = min(STAGE_R, a.ny - y0);
  const int inst = blockIdx.z;
  const InstParams P = a.ip[inst];
  const StageCoef<T> k{(T)P.alpha_a, (T)P.alpha_b, (T)P.F, (T)P.k, (T)P.dt, (T)(1.0 / a.dy2), (T)(1.0 / a.dx2)};
  if (xs + 64 * VE <= a.nx)
    stage_rows<T, METHOD, STAGE, VE, SAME, true>(a, k, inst, y0, rows, xs, lane);
  else
    stage_rows<T, METHOD, STAGE, VE, SAME, false>(a, k, inst, y0, rows, xs, lane);
}

template <typename T, int METHOD, int STAGE, int VE>
int stage_launch_ve(const Stage<T>& s, int B, hipStream_t st) {
  const bool same = s.zu == s.yu && s.zv == s.yv && s.zp == s.yp && s.zi == s.yi;
  const dim3 grid((unsigned)((s.nx + 64 * VE - 1) / (64 * VE)),
                  (unsigned)((s.ny + STAGE_WAVES * STAGE_R - 1) / (STAGE_WAVES * STAGE_R)), (unsigned)B);
  if (same)
    hipLaunchKernelGGL((k_gs_stage<T, METHOD, STAGE, VE, true>), grid, dim3(256), 0, st, s);
  else
    hipLaunchKernelGGL((k_gs_stage<T, METHOD, STAGE, VE, false>), grid, dim3(256), 0, st, s);
  return check_launch("gs_explicit_stage");
}

template <typename T, int METHOD, int STAGE>
int stage_launch(const Stage<T>& s, int B, hipStream_t st) {
  // columns per lane: one strip of 64 VE columns covers rows up to 256 wide
  if (s.nx <= 64) return stage_launch_ve<T, METHOD, STAGE, 1>(s, B, st);
  if (s.nx <= 128) return stage_launch_ve<T, METHOD, STAGE, 2>(s, B, st);
  return stage_launch_ve<T, METHOD, STAGE, 4>(s, B, st);
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
  // block: 256 consecutive columns of rows blockIdx.y, blockIdx.y + gridDim.y, ...
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nx) return;
  const int inst = blockIdx.z;
  for (int y = blockIdx.y; y < ny; y += gridDim.y) {
    const long long so = inst * src.istride + (long long)y * src.pitch + x;
    const long long d = inst * dst.istride + (long long)y * dst.pitch + x;
    dst.u[d] = src.u[so];
    dst.v[d] = src.v[so];
  }
}

// n_steps explicit steps in place on (u, v).  SSPRK3 and RK4 write each step
// straight back into (u, v): their stages before the last write scratch, and
// the last stage reads the step's base state only at the point it writes.
// Euler's one stage reads the stencil of the base state, so its steps
// alternate between two dense buffers; from the second step on the last
// step reads a dense buffer and writes (u, v), and a single step is copied
// back.
template <typename T>
int explicit_inplace_t(int method, int B, int ny, int nx, T* u, T* v, long long pitch, long long istride,
                       int ghosts, double dy, double dx, const InstParams* ip, int n_steps, T* work, hipStream_t st) {
  const long long n = (long long)ny * nx;
  T* scratch = work;
  T* P = work + explicit_scratch_elems(B, ny, nx);
  T* Q = P + (long long)B * 2 * n;
  const Layout2<T> state{u, v, pitch, istride};
  int rc = MJX_OK;
  if (method != EXPL_EULER) {
    for (int s = 0; s < n_steps && !rc; ++s)
      rc = explicit_step_t<T>(method, B, ny, nx, state, state, dy, dx, ip, scratch, st);
  } else {
    Layout2<T> cur = state;
    for (int s = 0; s < n_steps && !rc; ++s) {
      T* d = (s & 1) ? Q : P;
      const Layout2<T> out = (s > 0 && s == n_steps - 1) ? state : Layout2<T>{d, d + n, nx, 2 * n};
      rc = explicit_step_t<T>(method, B, ny, nx, cur, out, dy, dx, ip, scratch, st);
      cur = out;
    }
    if (!rc && n_steps == 1) {
      hipLaunchKernelGGL((k_copy2<T>), dim3((unsigned)((nx + 255) / 256), (unsigned)std::min(ny, 65535), B),
                         dim3(256), 0, st, cur, state, ny, nx);
      rc = check_launch("gs_explicit_copy");
    }
  }
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
  if (ny > 65535LL * 64) return fail(MJX_E_LIMIT, "at most 4194240 rows");
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
