// float instantiation of the batched Gray-Scott ETDRK4 chain.
#include "etd_kernels.h"

namespace mjx {

namespace {
using T = float;

template <int N>
int run_sized(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts, const T* coef,
              const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
  constexpr int LR = lines_default<T, N>() < N ? lines_default<T, N>() : N;           // rows per block
  constexpr int LC = lines_default<T, N>() < N / 2 ? lines_default<T, N>() : N / 2;   // columns per block
  const long long plane = (long long)B * N * N;
  const long long splane = (long long)B * N * (N + 2);
  cpx<T>* Ws = work;
  cpx<T>* Wg = Ws + plane;
  cpx<T>* Wn = Wg + plane;
  cpx<T>* Uh = Wn + plane;
  cpx<T>* Nn = Uh + splane;
  cpx<T>* Fh = Nn + splane;
  using E = EtdLaunch<T>;
  int rc = Launch<T>::row_fwd_pack(nx, B, ny, u, v, pitch, istride, Ws, st, (T)(1.0 / nx));
  for (int s = 0; s < n_steps && rc == 0; ++s) {
    rc = E::template row_nl<N, LR>(B, ny, Ws, Wn, ip, st);
    if (!rc) rc = E::template col<N, LC, 1>(B, nx, Ws, Wn, Uh, Nn, Fh, coef, Wg, st);
    if (!rc) rc = E::template row_nl<N, LR>(B, ny, Wg, Wn, ip, st);
    if (!rc) rc = E::template col<N, LC, 2>(B, nx, Wn, nullptr, Uh, Nn, Fh, coef, Wg, st);
    if (!rc) rc = E::template row_nl<N, LR>(B, ny, Wg, Wn, ip, st);
    if (!rc) rc = E::template col<N, LC, 3>(B, nx, Wn, nullptr, Uh, Nn, Fh, coef, Wg, st);
    if (!rc) rc = E::template row_nl<N, LR>(B, ny, Wg, Wn, ip, st);
    if (!rc) rc = E::template col<N, LC, 4>(B, nx, Wn, nullptr, Uh, Nn, Fh, coef, Ws, st);
  }
  if (rc) return rc;
  return Launch<T>::row_inv_unpack(nx, B, ny, Ws, u, v, pitch, istride, ghosts, st);
}
}  // namespace

int etd_run(int B, int n, T* u, T* v, long long pitch, long long istride, int ghosts, const T* coef,
            const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st) {
  switch (n) {
    case 32: return run_sized<32>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    case 64: return run_sized<64>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    case 128: return run_sized<128>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    case 256: return run_sized<256>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    case 512: return run_sized<512>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    case 1024: return run_sized<1024>(B, n, n, u, v, pitch, istride, ghosts, coef, ip, n_steps, work, st);
    default: return fail(MJX_E_SIZE, "ETDRK4 grid " + std::to_string(n) + " not supported: square n in 32..1024");
  }
}

int etd_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt, T* coef,
               hipStream_t st) {
  const long long n = (long long)ny * (nx + 2);
  hipLaunchKernelGGL(k_etd_coeffs<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lamx, lamy, ny, nx, Du, Dv,
                     dt, coef);
  return check_launch("etd_coeffs");
}

}  // namespace mjx
