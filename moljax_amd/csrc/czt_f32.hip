// FP32 instantiation of the chirp-z launchers.
#include "czt_launch.h"

namespace mjx {

using T = float;

int czt_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride, cpx<T>* W,
                     hipStream_t st, double in_scale) {
  return CztLaunch<T>::row_fwd_pack(nx, B, ny, a, b, pitch, istride, W, st, in_scale);
}
int czt_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch, long long istride,
                       int ghosts, hipStream_t st, double keep, const InstParams* react) {
  return CztLaunch<T>::row_inv_unpack(nx, B, ny, W, a, b, pitch, istride, ghosts, st, keep, react);
}
int czt_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep) {
  return CztLaunch<T>::row_gs(nx, B, ny, W, ip, st, rxn, mstep);
}
int czt_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,
                  const InstParams* ip, int nrep, hipStream_t st, int mstep) {
  return CztLaunch<T>::col_apply(ny, B, nx, W, lamx, lamy, lam2d, ip, nrep, st, mstep);
}
int czt_etdrk4_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts, const T* coef,
                   const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
  return CztLaunch<T>::etdrk4_run(B, ny, nx, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
}
int czt_etd_exp_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                    const T* coef, const InstParams* ip, int order, int prev, const T* nu, const T* nv, int n_steps,
                    cpx<T>* work, hipStream_t st) {
  return CztLaunch<T>::etd_exp_run(B, ny, nx, u, v, pitch, istride, ghosts, coef, ip, order, prev, nu, nv, n_steps,
                                   work, st);
}
int czt_etdrk4_pack(int B, int ny, int nx, const T* u, const T* v, long long pitch, long long istride, cpx<T>* work,
                    hipStream_t st) {
  return CztLaunch<T>::etdrk4_pack(B, ny, nx, u, v, pitch, istride, work, st);
}
int czt_etdrk4_advance(int B, int ny, int nx, const T* coef, const InstParams* ip, int n_steps, cpx<T>* work,
                       hipStream_t st) {
  return CztLaunch<T>::etdrk4_advance(B, ny, nx, coef, ip, n_steps, work, st);
}
int czt_etdrk4_unpack(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,
                      cpx<T>* work, hipStream_t st) {
  return CztLaunch<T>::etdrk4_unpack(B, ny, nx, u, v, pitch, istride, ghosts, work, st);
}
int czt_r2r_launch(const CztR2RArgs& a, int orient, bool apply, const CztTables& h1, const CztTables& h2, int B,
                   hipStream_t st, T) {
  return CztLaunch<T>::r2r(a, orient, apply, h1, h2, B, st);
}

}  // namespace mjx
