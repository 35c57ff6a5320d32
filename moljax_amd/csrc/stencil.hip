// Finite-difference operators, ghost-cell fill and point-wise reaction
// kernels on ghost-padded 2-D arrays.
//
// Arrays are addressed through a pointer to the first interior element, a row
// pitch and an instance stride (elements); the ghost ring sits at row/column
// -1 and ny/nx.  Output ghost cells of the FD operators are written as zero,
// as the reference does (moljax/core/operators.py:159-322 fill zeros_like and
// set the interior).
//
// Memory-bound streaming kernels with coalesced row access; the FD
// operators stage a tile and its halo in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>

#include "mjx_common.h"

namespace mjx {

enum FdOp { FD_LAP = 0, FD_D2 = 1, FD_D1C = 2, FD_D1UP = 3 };

// Batched stencil launches put the instance on grid y (grid z for the FD
// kernel), at most this many per launch; larger batches loop on the host.
constexpr long long MAX_GRID_Y = 65535;
constexpr int FD_TX = 64;  // padded columns per FD tile (one per lane)
constexpr int FD_TY = 32;  // padded rows per FD tile (8 per thread row)

struct FdArgs {
  long long ip, is;  // input pitch / instance stride
  long long op, os;  // output pitch / instance stride
  long long vp, vs;  // velocity pitch / instance stride (array velocity)
  int ny, nx, axis;
  double h, h2;      // spacing along axis and its square (D2 / D1 / upwind)
  double dx2, dy2;   // Laplacian squares
  double vscalar;
};

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_fd2d(const T* __restrict__ in, T* __restrict__ out,
                                              const T* __restrict__ vel, FdArgs a, long long b0) {
  // One 32 x 64 tile of the padded output per workgroup (64 x 4 threads).
  // The tile and its one-cell halo are staged in LDS with coalesced row
  // loads (each input value is read from memory once per tile, 10 % halo);
  // every thread then forms 8 outputs of one column from LDS and writes
  // them as coalesced row segments.  Output ghost cells are written as 0.
  __shared__ T tile[FD_TY + 2][FD_TX + 2];
  const int W = a.nx + 2, H = a.ny + 2;
  const int x0 = blockIdx.x * FD_TX, y0 = blockIdx.y * FD_TY;  // padded origin of the tile
  const long long b = b0 + blockIdx.z;
  const T* pin = in + b * a.is - a.ip - 1;  // padded (0, 0)
  const int tid = threadIdx.y * FD_TX + threadIdx.x;
  for (int idx = tid; idx < (FD_TY + 2) * (FD_TX + 2); idx += FD_TX * 4) {
    const int ly = idx / (FD_TX + 2), lx = idx - ly * (FD_TX + 2);
    const int yp = y0 - 1 + ly, xp = x0 - 1 + lx;
    T v = (T)0;
    if (yp >= 0 && yp < H && xp >= 0 && xp < W) v = pin[(long long)yp * a.ip + xp];
    tile[ly][lx] = v;
  }
  __syncthreads();
  const int lx = threadIdx.x, xp = x0 + lx;
  if (xp >= W) return;
  const int x = xp - 1;
  const bool xin = x >= 0 && x < a.nx;
  T* pout = out + b * a.os - a.op - 1;
#pragma unroll
  for (int r = 0; r < FD_TY / 4; ++r) {
    const int ly = threadIdx.y * (FD_TY / 4) + r;
    const int yp = y0 + ly;
    if (yp >= H) break;
    const int y = yp - 1;
    T res = (T)0;
    if (xin && y >= 0 && y < a.ny) {
      const T c = tile[ly + 1][lx + 1];
      const T up = tile[ly][lx + 1], dn = tile[ly + 2][lx + 1];
      const T lf = tile[ly + 1][lx], rt = tile[ly + 1][lx + 2];
      if constexpr (OP == FD_LAP) {
        // d2_central_2d(axis=0) + d2_central_2d(axis=1), operators.py:311-322
        res = (dn - (T)2 * c + up) / (T)a.dy2 + (rt - (T)2 * c + lf) / (T)a.dx2;
      } else {
        const T lo = a.axis == 0 ? up : lf;
        const T hi = a.axis == 0 ? dn : rt;
        if constexpr (OP == FD_D2) {
          res = (hi - (T)2 * c + lo) / (T)a.h2;  // operators.py:271-308
        } else if constexpr (OP == FD_D1C) {
          res = (hi - lo) / ((T)2 * (T)a.h);  // operators.py:159-197
        } else {
          // operators.py:200-268: backward where v >= 0, forward otherwise
          const T v = vel ? vel[b * a.vs + (long long)y * a.vp + x] : (T)a.vscalar;
          res = v >= (T)0 ? (c - lo) / (T)a.h : (hi - c) / (T)a.h;
        }
      }
    }
    pout[(long long)yp * a.op + xp] = res;
  }
}

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

template <typename T, int OP>
__device__ __forceinline__ T fd_point(T c, T up, T dn, T lf, T rt, const T* vel, const FdArgs& a, long long b, int y,
                                      int x) {
  if constexpr (OP == FD_LAP) {
    return (dn - (T)2 * c + up) / (T)a.dy2 + (rt - (T)2 * c + lf) / (T)a.dx2;  // operators.py:311-322
  } else {
    const T lo = a.axis == 0 ? up : lf;
    const T hi = a.axis == 0 ? dn : rt;
    if constexpr (OP == FD_D2) {
      return (hi - (T)2 * c + lo) / (T)a.h2;  // operators.py:271-308
    } else if constexpr (OP == FD_D1C) {
      return (hi - lo) / ((T)2 * (T)a.h);  // operators.py:159-197
    } else {
      // operators.py:200-268: backward where v >= 0, forward otherwise
      const T v = vel ? vel[b * a.vs + (long long)y * a.vp + x] : (T)a.vscalar;
      return v >= (T)0 ? (c - lo) / (T)a.h : (hi - c) / (T)a.h;
    }
  }
}

// The FD tile kernel with two-element (8-byte FP32, 16-byte FP64) accesses,
// for arrays whose padded rows start on a two-element boundary (even pitch
// and instance stride, aligned origin): the LDS tile holds padded columns
// [x0 - 2, x0 + 66) as 34 pairs per row; each thread forms a pair of
// adjacent outputs in 4 rows and stores it as one pair.
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_fd2d_pair(const T* __restrict__ in, T* __restrict__ out,
                                                   const T* __restrict__ vel, FdArgs a, long long b0) {
  using P = typename Vec2<T>::type;
  constexpr int NP = FD_TX / 2 + 2;  // pairs per LDS row
  __shared__ P tile[FD_TY + 2][NP];
  const int W = a.nx + 2, H = a.ny + 2;
  const int x0 = blockIdx.x * FD_TX, y0 = blockIdx.y * FD_TY;
  const long long b = b0 + blockIdx.z;
  const T* pin = in + b * a.is - a.ip - 1;  // padded (0, 0), two-element aligned
  const int tid = threadIdx.y * FD_TX + threadIdx.x;
  for (int idx = tid; idx < (FD_TY + 2) * NP; idx += FD_TX * 4) {
    const int ly = idx / NP, q = idx - ly * NP;
    const int yp = y0 - 1 + ly, c = x0 - 2 + 2 * q;
    P v;
    v.x = (T)0;
    v.y = (T)0;
    if (yp >= 0 && yp < H) {
      const T* row = pin + (long long)yp * a.ip;
      if (c >= 0 && c + 1 < W) {
        v = *reinterpret_cast<const P*>(row + c);
      } else {
        if (c >= 0 && c < W) v.x = row[c];
        if (c + 1 >= 0 && c + 1 < W) v.y = row[c + 1];
      }
    }
    tile[ly][q] = v;
  }
  __syncthreads();
  const int cp = tid & 31, rg = tid >> 5;
  const int xp0 = x0 + 2 * cp;
  if (xp0 >= W) return;
  T* pout = out + b * a.os - a.op - 1;
  const int xa = xp0 - 1, xb = xp0;  // interior x of the two outputs
  const bool ina = xa >= 0 && xa < a.nx, inb = xb >= 0 && xb < a.nx;
#pragma unroll
  for (int r = 0; r < FD_TY / 8; ++r) {
    const int ly = rg * (FD_TY / 8) + r;
    const int yp = y0 + ly;
    if (yp >= H) break;
    const int y = yp - 1;
    const bool yin = y >= 0 && y < a.ny;
    const P C = tile[ly + 1][cp + 1], U = tile[ly][cp + 1], D = tile[ly + 2][cp + 1];
    const P L = tile[ly + 1][cp], R = tile[ly + 1][cp + 2];
    P o;
    o.x = (yin && ina) ? fd_point<T, OP>(C.x, U.x, D.x, L.y, C.y, vel, a, b, y, xa) : (T)0;
    o.y = (yin && inb) ? fd_point<T, OP>(C.y, U.y, D.y, C.x, R.x, vel, a, b, y, xb) : (T)0;
    T* orow = pout + (long long)yp * a.op;
    if (xp0 + 1 < W) {
      *reinterpret_cast<P*>(orow + xp0) = o;
    } else {
      orow[xp0] = o.x;
    }
  }
}

// The FD operators on instances whose padded rows are contiguous (pitch =
// nx + 2, even) with 16-byte aligned origins: the instance is one flat
// stream of (ny + 2)(nx + 2) values, each thread forms VE consecutive
// outputs from one 16-byte load of the centres, the 16-byte-pair loads of
// the values one row above and below, and the two end neighbours, and
// writes them with one 16-byte store.  The rows above and below were loaded
// by neighbouring workgroups just before, so they come from cache; HBM sees
// each value read once and written once.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double2;
};

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_fd2d_flat(const T* __restrict__ in, T* __restrict__ out,
                                                   const T* __restrict__ vel, FdArgs a, long long b0, double invW) {
  constexpr int VE = 16 / sizeof(T);  // outputs per thread
  using V = typename Vec4<T>::type;
  using P = typename Vec2<T>::type;
  const int W = a.nx + 2, H = a.ny + 2;
  const int N = H * W;
  const int i0 = (blockIdx.x * 256 + threadIdx.x) * VE;
  if (i0 >= N) return;
  const long long b = b0 + blockIdx.y;
  const T* p = in + b * a.is - a.ip - 1;  // padded (0, 0)
  T* q = out + b * a.os - a.op - 1;
  T c[VE], up[VE], dn[VE];
  {
    const V cv = *reinterpret_cast<const V*>(p + i0);
    const T* cs = reinterpret_cast<const T*>(&cv);
#pragma unroll
    for (int e = 0; e < VE; ++e) c[e] = cs[e];
  }
  if (i0 >= W) {
#pragma unroll
    for (int e = 0; e < VE; e += 2) {
      const P v = *reinterpret_cast<const P*>(p + i0 - W + e);
      up[e] = v.x;
      up[e + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < VE; ++e) up[e] = i0 + e >= W ? p[i0 + e - W] : (T)0;
  }
  if (i0 + VE - 1 + W < N) {
#pragma unroll
    for (int e = 0; e < VE; e += 2) {
      const P v = *reinterpret_cast<const P*>(p + i0 + W + e);
      dn[e] = v.x;
      dn[e + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < VE; ++e) dn[e] = i0 + e + W < N ? p[i0 + e + W] : (T)0;
  }
  const T lf0 = i0 > 0 ? p[i0 - 1] : (T)0;
  const T rtN = i0 + VE < N ? p[i0 + VE] : (T)0;
  // row and column of the first output (exact: the quotient is corrected)
  int row = (int)((double)i0 * invW);
  int col = i0 - row * W;
  if (col < 0) {
    --row;
    col += W;
  } else if (col >= W) {
    ++row;
    col -= W;
  }
  T o[VE];
#pragma unroll
  for (int e = 0; e < VE; ++e) {
    const int y = row - 1, x = col - 1;
    const T lf = e == 0 ? lf0 : c[e - 1];
    const T rt = e == VE - 1 ? rtN : c[e + 1];
    o[e] = (y >= 0 && y < a.ny && x >= 0 && x < a.nx) ? fd_point<T, OP>(c[e], up[e], dn[e], lf, rt, vel, a, b, y, x)
                                                       : (T)0;
    if (++col == W) {
      col = 0;
      ++row;
    }
  }
  V ov;
  T* os = reinterpret_cast<T*>(&ov);
#pragma unroll
  for (int e = 0; e < VE; ++e) os[e] = o[e];
  *reinterpret_cast<V*>(q + i0) = ov;
}

// --------------------------------------------------------------- BC fill
// kind: 0 periodic, 1 Dirichlet (ghost = 2 g - interior), 2 Neumann
// (ghost = interior -/+ 2 h flux); side values [left, right, bottom, top].
// Final ghost values are computed directly from interior values in the order
// of moljax/core/bc.py:202-307 (x ghosts of every row first, then full ghost
// rows including the x ghosts), so corners follow the reference.
struct BcArgs {
  long long p, is;
  int ny, nx, kind;
  double vl, vr, vb, vt;
  double dx, dy;
};

template <typename T>
__device__ __forceinline__ T xghost(const T* row, int side, const BcArgs& a) {
  // row: pointer to interior col 0 of an interior row
  if (a.kind == 0) return side == 0 ? row[a.nx - 1] : row[0];
  if (a.kind == 1) return side == 0 ? (T)(2.0 * a.vl) - row[0] : (T)(2.0 * a.vr) - row[a.nx - 1];
  return side == 0 ? row[0] - (T)(2.0 * a.dx * a.vl) : row[a.nx - 1] + (T)(2.0 * a.dx * a.vr);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_fill(T* __restrict__ f, BcArgs a, int per, long long b0) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;  // instance on grid y
  if (r >= per) return;
  const long long b = b0 + blockIdx.y;
  T* base = f + b * a.is;
  const int W = a.nx + 2;
  if (r < 2 * a.ny) {
    // x ghosts of interior rows: r = 2*y + side
    const int y = r >> 1, side = r & 1;
    T* row = base + (long long)y * a.p;
    const T val = xghost(row, side, a);
    row[side == 0 ? -1 : a.nx] = val;
    return;
  }
  r -= 2 * a.ny;
  // ghost rows: r = side*W + (c+1), c in [-1, nx]
  const int side = r / W, c = r % W - 1;
  int src;
  if (a.kind == 0) src = side == 0 ? a.ny - 1 : 0;
  else src = side == 0 ? 0 : a.ny - 1;
  const T* srow = base + (long long)src * a.p;
  T s;
  if (c < 0) s = xghost(srow, 0, a);
  else if (c >= a.nx) s = xghost(srow, 1, a);
  else s = srow[c];
  T val;
  if (a.kind == 0) val = s;
  else if (a.kind == 1) val = side == 0 ? (T)(2.0 * a.vb) - s : (T)(2.0 * a.vt) - s;
  else val = side == 0 ? s - (T)(2.0 * a.dy * a.vb) : s + (T)(2.0 * a.dy * a.vt);
  base[(side == 0 ? -1LL : (long long)a.ny) * a.p + c] = val;
}

// ------------------------------------------------------------ reaction
// Gray-Scott reaction on whole (padded) arrays, operators.py:655-667.
template <typename T>
__global__ __launch_bounds__(256) void k_gs_reaction(const T* __restrict__ u, const T* __restrict__ v,
                                                     T* __restrict__ ru, T* __restrict__ rv, long long n, double F,
                                                     double k) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const T uu = u[g], vv = v[g];
  const T uvv = uu * vv * vv;
  ru[g] = -uvv + (T)F * ((T)1 - uu);
  rv[g] = uvv - ((T)F + (T)k) * vv;
}

static inline unsigned nblocks(long long total) { return (unsigned)((total + 255) / 256); }


// out = a x + b y + c over n contiguous elements (y may be null): the
// element-wise StateDict algebra (moljax/core/state.py:27-81).
template <typename T>
__global__ __launch_bounds__(256) void k_axpby(long long n, double a, const T* x, double b, const T* y, double c,
                                               T* out) {  // out may alias x or y: no restrict
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T v = (T)a * x[i];
  if (y) v = v + (T)b * y[i];
  if (c != 0.0) v = v + (T)c;
  out[i] = v;
}

// One weighted-Jacobi sweep for (I - beta lap) x = r on the whole padded
// array (moljax/core/preconditioners.py:166-198): x' = omega (r + beta
// lap(x)) / diag + (1 - omega) x with lap the 5-point Laplacian on interior
// points and 0 on the ghost ring (as laplacian_2d writes it).
struct JacobiArgs {
  long long rp, rs, xp, xs, op, os;  // pitches / instance strides of r, x, out
  int ny, nx;
  double beta, dx2, dy2, omega, diag;
};

template <typename T>
__global__ __launch_bounds__(256) void k_jacobi_padded(const T* __restrict__ r, const T* __restrict__ x,
                                                       T* __restrict__ out, JacobiArgs a, int per, long long b0) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;  // instance on grid y
  if (rr >= per) return;
  const long long b = b0 + blockIdx.y;
  const int W = a.nx + 2;
  const int row = rr / W;
  const int y = row - 1, xi = rr - row * W - 1;
  const T* xc = x + b * a.xs + (long long)y * a.xp + xi;
  const T c = *xc;
  T lap = (T)0;
  if (y >= 0 && y < a.ny && xi >= 0 && xi < a.nx)
    lap = (xc[a.xp] - (T)2 * c + xc[-a.xp]) / (T)a.dy2 + (xc[1] - (T)2 * c + xc[-1]) / (T)a.dx2;
  const T rv = r[b * a.rs + (long long)y * a.rp + xi];
  out[b * a.os + (long long)y * a.op + xi] =
      (T)a.omega * (rv + (T)a.beta * lap) / (T)a.diag + ((T)1 - (T)a.omega) * c;
}

}  // namespace mjx

using namespace mjx;

extern "C" {

int mjx_axpby(int dtype, long long n, double a, const void* x, double b, const void* y, double c, void* out,
              void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (n < 0 || (n > 0 && (!x || !out))) return fail(MJX_E_ARG, "null pointer or negative length");
  if (n == 0) return MJX_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MJX_F32)
    hipLaunchKernelGGL(k_axpby<float>, dim3(nblocks(n)), dim3(256), 0, st, n, a, (const float*)x, b,
                       (const float*)y, c, (float*)out);
  else
    hipLaunchKernelGGL(k_axpby<double>, dim3(nblocks(n)), dim3(256), 0, st, n, a, (const double*)x, b,
                       (const double*)y, c, (double*)out);
  return check_launch("axpby");
}

int mjx_jacobi2d(int dtype, int B, int ny, int nx, const void* r, long long r_pitch, long long r_istride, void* x,
                 long long x_pitch, long long x_istride, void* work, double beta, double dy, double dx, double omega,
                 int n_iter, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (B < 1 || ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty grid");
  if (!r || !x || (n_iter > 1 && !work)) return fail(MJX_E_ARG, "null pointer argument");
  if (n_iter < 0) return fail(MJX_E_ARG, "n_iter must be >= 0");
  if (r_pitch < nx + 2 || x_pitch < nx + 2) return fail(MJX_E_ARG, "pitch must cover the ghost ring");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(MJX_E_ARG, "dx, dy must be positive");
  if ((long long)(ny + 2) * (nx + 2) > INT_MAX) return fail(MJX_E_LIMIT, "grid too large for one instance");
  const long long W = nx + 2;
  const int per = (ny + 2) * (nx + 2);
  const size_t es = dtype == MJX_F64 ? 8 : 4;
  JacobiArgs a;
  a.ny = ny;
  a.nx = nx;
  a.beta = beta;
  a.dx2 = dx * dx;
  a.dy2 = dy * dy;
  a.diag = 1.0 + beta * (2.0 / (dx * dx) + 2.0 / (dy * dy));
  hipStream_t st = (hipStream_t)stream;
  // x0 = r; sweeps alternate between x and work so that the last lands in x
  // (n_iter = 0: one sweep with omega 0, a copy)
  const int sweeps = n_iter > 0 ? n_iter : 1;
  const char* prev = (const char*)r;
  long long pp = r_pitch, ps = r_istride;
  char* wk = work ? (char*)work + es * (W + 1) : nullptr;  // origin of the dense padded work array
  for (int i = 0; i < sweeps; ++i) {
    const bool to_x = (sweeps - 1 - i) % 2 == 0;
    char* dst = to_x ? (char*)x : wk;
    a.rp = r_pitch;
    a.rs = r_istride;
    a.xp = pp;
    a.xs = ps;
    a.op = to_x ? x_pitch : W;
    a.os = to_x ? x_istride : per;
    a.omega = n_iter > 0 ? omega : 0.0;
    for (long long b0 = 0; b0 < B; b0 += MAX_GRID_Y) {
      const dim3 grid((per + 255) / 256, (unsigned)std::min<long long>(MAX_GRID_Y, B - b0));
      if (dtype == MJX_F32)
        hipLaunchKernelGGL(k_jacobi_padded<float>, grid, dim3(256), 0, st, (const float*)r, (const float*)prev,
                           (float*)dst, a, per, b0);
      else
        hipLaunchKernelGGL(k_jacobi_padded<double>, grid, dim3(256), 0, st, (const double*)r,
                           (const double*)prev, (double*)dst, a, per, b0);
    }
    int rc = check_launch("jacobi_padded");
    if (rc) return rc;
    prev = dst;
    pp = a.op;
    ps = a.os;
  }
  return MJX_OK;
}

int mjx_fd_op2d(int dtype, int op, int B, int ny, int nx, const void* in, long long in_pitch, long long in_istride,
                void* out, long long out_pitch, long long out_istride, int axis, double dx, double dy,
                const void* vel, long long vel_pitch, long long vel_istride, double vel_scalar, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (op < FD_LAP || op > FD_D1UP) return fail(MJX_E_ARG, "unknown FD operator");
  if (B < 1 || ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty grid");
  if (!in || !out) return fail(MJX_E_ARG, "null pointer argument");
  if (axis != 0 && axis != 1) return fail(MJX_E_ARG, "axis must be 0 (y) or 1 (x)");
  if (in_pitch < nx + 2 || out_pitch < nx + 2) return fail(MJX_E_ARG, "pitch must cover the ghost ring");
  if (vel && op == FD_D1UP && (vel_pitch < nx + 2 || (B > 1 && vel_istride < vel_pitch * (ny + 2))))
    return fail(MJX_E_ARG, "velocity pitch must cover the ghost ring and instances must not overlap");
  FdArgs a;
  a.ip = in_pitch;
  a.is = in_istride;
  a.op = out_pitch;
  a.os = out_istride;
  a.vp = vel_pitch;
  a.vs = vel_istride;
  a.ny = ny;
  a.nx = nx;
  a.axis = axis;
  a.h = axis == 0 ? dy : dx;
  a.h2 = a.h * a.h;
  a.dx2 = dx * dx;
  a.dy2 = dy * dy;
  a.vscalar = vel_scalar;
  if (nx > INT_MAX - 66 || ny > 65535 * FD_TY - 2) return fail(MJX_E_LIMIT, "grid too large for one launch");
  hipStream_t st = (hipStream_t)stream;
  // two-element accesses when every padded row starts on a two-element
  // boundary of input, output and velocity
  const size_t es = dtype == MJX_F64 ? 8 : 4;
  auto pair_ok = [&](const void* p, long long pitch, long long istride) {
    if (!p) return true;
    const uintptr_t origin = (uintptr_t)p - (uintptr_t)((pitch + 1) * (long long)es);
    return pitch % 2 == 0 && (B == 1 || istride % 2 == 0) && origin % (2 * es) == 0;
  };
  const bool pairs = pair_ok(in, in_pitch, in_istride) && pair_ok(out, out_pitch, out_istride);
  // the flat form: contiguous padded rows (pitch = nx + 2, even), 16-byte
  // aligned instance origins, a whole number of 16-byte units per instance
  const long long Wp = nx + 2, Np = (long long)(ny + 2) * Wp;
  const long long ve = 16 / (long long)es;
  auto flat_ok = [&](const void* ptr, long long pitch, long long istride) {
    const uintptr_t origin = (uintptr_t)ptr - (uintptr_t)((pitch + 1) * (long long)es);
    return pitch == Wp && Wp % 2 == 0 && Np % ve == 0 && (B == 1 || istride % ve == 0) && origin % 16 == 0;
  };
  const bool flat = Np < (1LL << 30) && flat_ok(in, in_pitch, in_istride) && flat_ok(out, out_pitch, out_istride);
  if (flat) {
    const double invW = 1.0 / (double)Wp;
    const unsigned gx = (unsigned)((Np / ve + 255) / 256);
#define MJX_FD_FLAT(T, OPV)                                                                           \
  for (long long b0 = 0; b0 < B; b0 += MAX_GRID_Y)                                                    \
  hipLaunchKernelGGL((k_fd2d_flat<T, OPV>), dim3(gx, (unsigned)std::min<long long>(MAX_GRID_Y, B - b0)), \
                     dim3(256), 0, st, (const T*)in, (T*)out, (const T*)vel, a, b0, invW)
    if (dtype == MJX_F32) {
      switch (op) {
        case FD_LAP: MJX_FD_FLAT(float, FD_LAP); break;
        case FD_D2: MJX_FD_FLAT(float, FD_D2); break;
        case FD_D1C: MJX_FD_FLAT(float, FD_D1C); break;
        default: MJX_FD_FLAT(float, FD_D1UP); break;
      }
    } else {
      switch (op) {
        case FD_LAP: MJX_FD_FLAT(double, FD_LAP); break;
        case FD_D2: MJX_FD_FLAT(double, FD_D2); break;
        case FD_D1C: MJX_FD_FLAT(double, FD_D1C); break;
        default: MJX_FD_FLAT(double, FD_D1UP); break;
      }
    }
#undef MJX_FD_FLAT
    return check_launch("fd_op2d");
  }
#define MJX_FD_LAUNCH(T, OPV)                                                                         \
  for (long long b0 = 0; b0 < B; b0 += MAX_GRID_Y)                                                    \
  hipLaunchKernelGGL(pairs ? (k_fd2d_pair<T, OPV>) : (k_fd2d<T, OPV>),                                \
                     dim3((nx + 2 + FD_TX - 1) / FD_TX, (ny + 2 + FD_TY - 1) / FD_TY,                   \
                          (unsigned)std::min<long long>(MAX_GRID_Y, B - b0)),                         \
                     dim3(64, 4), 0, st, (const T*)in, (T*)out, (const T*)vel, a, b0)
  if (dtype == MJX_F32) {
    switch (op) {
      case FD_LAP: MJX_FD_LAUNCH(float, FD_LAP); break;
      case FD_D2: MJX_FD_LAUNCH(float, FD_D2); break;
      case FD_D1C: MJX_FD_LAUNCH(float, FD_D1C); break;
      default: MJX_FD_LAUNCH(float, FD_D1UP); break;
    }
  } else {
    switch (op) {
      case FD_LAP: MJX_FD_LAUNCH(double, FD_LAP); break;
      case FD_D2: MJX_FD_LAUNCH(double, FD_D2); break;
      case FD_D1C: MJX_FD_LAUNCH(double, FD_D1C); break;
      default: MJX_FD_LAUNCH(double, FD_D1UP); break;
    }
  }
#undef MJX_FD_LAUNCH
  return check_launch("fd_op2d");
}

int mjx_bc_fill2d(int dtype, int B, int ny, int nx, void* f, long long pitch, long long istride, int kind,
                  const double* side_values, double dx, double dy, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (B < 1 || ny < 1 || nx < 1) return fail(MJX_E_ARG, "empty grid");
  if (!f) return fail(MJX_E_ARG, "null pointer argument");
  if (kind < 0 || kind > 2) return fail(MJX_E_ARG, "kind must be 0 (periodic), 1 (Dirichlet) or 2 (Neumann)");
  if (pitch < nx + 2) return fail(MJX_E_ARG, "pitch must cover the ghost ring");
  BcArgs a;
  a.p = pitch;
  a.is = istride;
  a.ny = ny;
  a.nx = nx;
  a.kind = kind;
  a.vl = side_values ? side_values[0] : 0.0;
  a.vr = side_values ? side_values[1] : 0.0;
  a.vb = side_values ? side_values[2] : 0.0;
  a.vt = side_values ? side_values[3] : 0.0;
  a.dx = dx;
  a.dy = dy;
  if ((long long)ny + nx > INT_MAX / 4) return fail(MJX_E_LIMIT, "grid too large for one instance");
  const int per = 2 * ny + 2 * (nx + 2);
  hipStream_t st = (hipStream_t)stream;
  // x ghosts of the interior rows must be final before the ghost rows copy
  // them; the ghost rows compute them again from interior values instead, so
  // one launch suffices and no ordering between the two groups is needed.
  for (long long b0 = 0; b0 < B; b0 += MAX_GRID_Y) {
    const dim3 grid((per + 255) / 256, (unsigned)std::min<long long>(MAX_GRID_Y, B - b0));
    if (dtype == MJX_F32)
      hipLaunchKernelGGL(k_bc_fill<float>, grid, dim3(256), 0, st, (float*)f, a, per, b0);
    else
      hipLaunchKernelGGL(k_bc_fill<double>, grid, dim3(256), 0, st, (double*)f, a, per, b0);
  }
  return check_launch("bc_fill2d");
}

int mjx_gs_reaction(int dtype, long long n, const void* u, const void* v, void* ru, void* rv, double F, double k,
                    void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (n < 0) return fail(MJX_E_ARG, "negative size");
  if (n == 0) return MJX_OK;
  if (!u || !v || !ru || !rv) return fail(MJX_E_ARG, "null pointer argument");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MJX_F32)
    hipLaunchKernelGGL(k_gs_reaction<float>, dim3(nblocks(n)), dim3(256), 0, st, (const float*)u, (const float*)v,
                       (float*)ru, (float*)rv, n, F, k);
  else
    hipLaunchKernelGGL(k_gs_reaction<double>, dim3(nblocks(n)), dim3(256), 0, st, (const double*)u,
                       (const double*)v, (double*)ru, (double*)rv, n, F, k);
  return check_launch("gs_reaction");
}

}  // extern "C"
