// FP32 instantiation of the periodic spectral launchers.
#include "czt_api.h"
#include "spectral_api.h"
#include "spectral_kernels.h"

namespace mjx {

using T = float;

int spec_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                      hipStream_t st, double in_scale) {
  if (spec_generic(nx, ny)) return czt_row_fwd_pack(nx, B, ny, a, b, pitch, istride, W, st, in_scale);
  return Launch<T>::row_fwd_pack(nx, B, ny, a, b, pitch, istride, W, st, (T)in_scale);
}
int spec_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                        int ghosts, hipStream_t st, double keep, const InstParams* react) {
  if (spec_generic(nx, ny)) return czt_row_inv_unpack(nx, B, ny, W, a, b, pitch, istride, ghosts, st, keep, react);
  return Launch<T>::row_inv_unpack(nx, B, ny, W, a, b, pitch, istride, ghosts, st, (T)keep, react);
}
int spec_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) {
  if (spec_generic(nx, ny)) return czt_row_gs(nx, B, ny, W, ip, st, rxn, mstep);
  return Launch<T>::row_gs(nx, B, ny, W, ip, st, rxn, mstep);
}
int spec_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                   const InstParams* ip, int nrep, hipStream_t st, int mstep) {
  if (spec_generic(nx, ny)) return czt_col_apply(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep);
  return Launch<T>::col_apply(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep);
}

}  // namespace mjx
