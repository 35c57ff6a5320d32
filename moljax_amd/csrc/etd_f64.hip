// double instantiation of the batched Gray-Scott ETDRK4 chain.
#include "etd_kernels.h"

namespace mjx {

using T = double;

int etd_run(int B, int n, T* u, T* v, long long pitch, long long istride, int ghosts, const T* coef,
            const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
  const EtdPlanes<T> p = etd_planes(work, B, n);
  int rc = Launch<T, SET_POW2>::row_fwd_pack(n, B, n, u, v, pitch, istride, p.Ws, st, (T)(1.0 / n));
  if (!rc) rc = etd_steps<T>(B, n, p, coef, ip, n_steps, st);
  if (rc) return rc;
  return Launch<T, SET_POW2>::row_inv_unpack(n, B, n, p.Ws, u, v, pitch, istride, ghosts, st);
}

int etd_pack(int B, int n, const T* u, const T* v, long long pitch, long long istride, cpx<T>* work, hipStream_t st) {
  return Launch<T, SET_POW2>::row_fwd_pack(n, B, n, u, v, pitch, istride, etd_planes(work, B, n).Ws, st, (T)(1.0 / n));
}

int etd_unpack(int B, int n, T* u, T* v, long long pitch, long long istride, int ghosts, cpx<T>* work,
               hipStream_t st) {
  return Launch<T, SET_POW2>::row_inv_unpack(n, B, n, etd_planes(work, B, n).Ws, u, v, pitch, istride, ghosts, st);
}

int etd_advance(int B, int n, const T* coef, const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
  return etd_steps<T>(B, n, etd_planes(work, B, n), coef, ip, n_steps, st);
}

// Exponential Euler (order 1) / AB2 (order 2).  prev: 0 = no previous
// nonlinear term (the first AB2 step is ETD1), 1 = the physical N_{n-1} in
// (nu, nv) with the state's layout, 2 = the workspace's N^ plane from the
// previous call on this workspace.
int etd_exp_run(int B, int n, T* u, T* v, long long pitch, long long istride, int ghosts, const T* coef,
                const InstParams* ip, int order, int prev, const T* nu, const T* nv, int n_steps, cpx<T>* work,
                hipStream_t st) {
  const EtdPlanes<T> p = etd_planes(work, B, n);
  int rc = MJX_OK;
  if (order == 2 && prev == 1) {
    rc = Launch<T, SET_POW2>::row_fwd_pack(n, B, n, nu, nv, pitch, istride, p.P1, st, (T)(1.0 / n));
    if (!rc) rc = etd_store_n<T>(B, n, p, p.P1, st);
  }
  if (!rc) rc = Launch<T, SET_POW2>::row_fwd_pack(n, B, n, u, v, pitch, istride, p.Ws, st, (T)(1.0 / n));
  cpx<T>* cur = p.Ws;
  if (!rc) rc = etd_exp_steps<T>(B, n, p, cur, coef, ip, order, prev != 0, n_steps, st);
  if (rc) return rc;
  return Launch<T, SET_POW2>::row_inv_unpack(n, B, n, cur, u, v, pitch, istride, ghosts, st);
}

int etd_phi_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt, T* coef,
                   hipStream_t st) {
  const long long n = (long long)ny * (nx + 2);
  hipLaunchKernelGGL(k_etd_phi_coeffs<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lamx, lamy, ny, nx,
                     Du, Dv, dt, coef);
  return check_launch("etd_phi_coeffs");
}

int etd_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt, T* coef,
               double f2w, hipStream_t st) {
  const long long n = (long long)ny * (nx + 2);
  hipLaunchKernelGGL(k_etd_coeffs<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lamx, lamy, ny, nx, Du, Dv,
                     dt, coef, f2w);
  return check_launch("etd_coeffs");
}

}  // namespace mjx
