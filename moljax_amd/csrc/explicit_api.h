// Explicit Runge-Kutta steps of the periodic Gray-Scott model (explicit.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mjx_common.h"

namespace mjx {

enum { EXPL_EULER = 0, EXPL_SSPRK3 = 1, EXPL_RK4 = 2 };

// Two fields of B instances: instance b's u at u + b*istride, rows pitch apart.
template <typename T>
struct Layout2 {
  T* u;
  T* v;
  long long pitch, istride;
};

// Scratch elements needed by gs_explicit_step: three dense two-field buffers.
inline long long explicit_scratch_elems(int B, int ny, int nx) { return 6LL * B * ny * nx; }

// One explicit step y -> out.  out must not overlap the scratch; for Euler it
// must not overlap y either, while SSPRK3 and RK4 accept out == y (same
// layout): only their last stage writes out, reading y at the written point.
int gs_explicit_step(int method, int B, int ny, int nx, const Layout2<float>& y, const Layout2<float>& out,
                     double dy, double dx, const InstParams* ip, float* scratch, hipStream_t st);
int gs_explicit_step(int method, int B, int ny, int nx, const Layout2<double>& y, const Layout2<double>& out,
                     double dy, double dx, const InstParams* ip, double* scratch, hipStream_t st);

}  // namespace mjx
