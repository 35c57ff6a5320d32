// FP64 instantiation of the periodic spectral launchers (power-of-two lengths).
#include "czt_api.h"
#include "spectral_api.h"
#include "spectral_kernels.h"

namespace mjx {

using T = double;

// The mixed-radix launchers are compiled in spectral_m5_f64.hip and
// spectral_m3_f64.hip.
extern template struct Launch<T, SET_MIX5>;
extern template struct Launch<T, SET_MIX3>;

// Launcher set of a line length (unsupported lengths fail in the set's switch).
#define MJX_BY_SET(n, CALL)                    \
  switch (size_set(n)) {                       \
    case SET_POW2:                             \
      return Launch<T, SET_POW2>::CALL;        \
    case SET_MIX5:                             \
      return Launch<T, SET_MIX5>::CALL;        \
    default:                                   \
      return Launch<T, SET_MIX3>::CALL;        \
  }

int spec_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                      hipStream_t st, double in_scale) {
  if (spec_generic(nx, ny)) return czt_row_fwd_pack(nx, B, ny, a, b, pitch, istride, W, st, in_scale);
  MJX_BY_SET(nx, row_fwd_pack(nx, B, ny, a, b, pitch, istride, W, st, (T)in_scale))
}
int spec_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                        int ghosts, hipStream_t st, double keep, const InstParams* react) {
  if (spec_generic(nx, ny)) return czt_row_inv_unpack(nx, B, ny, W, a, b, pitch, istride, ghosts, st, keep, react);
  MJX_BY_SET(nx, row_inv_unpack(nx, B, ny, W, a, b, pitch, istride, ghosts, st, (T)keep, react))
}
int spec_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) {
  if (spec_generic(nx, ny)) return czt_row_gs(nx, B, ny, W, ip, st, rxn, mstep);
  MJX_BY_SET(nx, row_gs(nx, B, ny, W, ip, st, rxn, mstep))
}
int spec_row_gs_inc(int nx, int B, int ny, cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                    const InstParams* ip, int mode, int ghosts, hipStream_t st) {
  if (!is_pow2_len(nx) || !is_pow2_len(ny))
    return fail(MJX_E_SIZE, "increment-form IMEX-Strang: nx and ny must be powers of two in [16, 4096]");
  return Launch<T, SET_POW2>::row_gs_inc(nx, B, ny, W, a, b, pitch, istride, ip, mode, ghosts, st);
}
int spec_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                   const InstParams* ip, int nrep, hipStream_t st, int mstep) {
  if (spec_generic(nx, ny)) return czt_col_apply(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep);
  MJX_BY_SET(ny, col_apply(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep))
}

}  // namespace mjx
