// Host-side launch API of the periodic spectral kernels, one overload per
// precision (defined in spectral_f32.hip / spectral_f64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mjx_common.h"
#include "mjx_fft.h"

namespace mjx {

// Explicit reaction stage of the fused row pass (k_row_gs): Heun over dt,
// forward Euler over dt, over dt/2, or two forward-Euler steps of dt/2.
enum { RXN_HEUN = 0, RXN_EULER = 1, RXN_HALF = 2, RXN_HALF2 = 3 };
// Column-pass power taken per instance from the instance row's aux column.
constexpr int NREP_PER_INSTANCE = -3;
// Column-pass mask: skip instances whose row has aux <= 0.
constexpr int MSTEP_ACTIVE = -2;
// Column-pass increment multipliers (runtime nrep): g - 1 and g^2 - 1.
constexpr int NREP_INC1 = -4;
constexpr int NREP_INC2 = -5;
// Row passes of the increment-form IMEX-Strang chain (k_row_gs_inc): the
// opening transform, a step's update + reaction + transform, the closing
// update.
enum { INC_FIRST = 0, INC_MID = 1, INC_LAST = 2 };

#define MJX_SPEC_DECL(T)                                                                                     \
  int spec_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,   \
                        cpx<T>* W, hipStream_t st, double in_scale = 1.0);                                   \
  int spec_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,              \
                          long long istride, int ghosts, hipStream_t st, double keep = 0.0,                  \
                          const InstParams* react = nullptr);                                                \
  int spec_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn = 0,      \
                  int mstep = -1);                                                                           \
  int spec_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,        \
                     const InstParams* ip, int nrep, hipStream_t st, int mstep = -1);                         \
  int spec_row_gs_inc(int nx, int B, int ny, cpx<T>* W, T* a, T* b, long long pitch, long long istride,      \
                      const InstParams* ip, int mode, int ghosts, hipStream_t st);

MJX_SPEC_DECL(float)
MJX_SPEC_DECL(double)
#undef MJX_SPEC_DECL

}  // namespace mjx
