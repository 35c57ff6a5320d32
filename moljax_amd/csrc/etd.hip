// C-ABI entry points of the batched Gray-Scott ETDRK4 path.
#include <hip/hip_runtime.h>

#include <string>

#include "czt_api.h"
#include "mjx_common.h"
#include "mjx_fft.h"

namespace mjx {
int etd_run(int B, int n, float* u, float* v, long long pitch, long long istride, int ghosts, const float* coef,
            const InstParams* ip, int n_steps, cpx<float>* work, hipStream_t st);
int etd_run(int B, int n, double* u, double* v, long long pitch, long long istride, int ghosts, const double* coef,
            const InstParams* ip, int n_steps, cpx<double>* work, hipStream_t st);
int etd_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt, float* coef,
               double f2w, hipStream_t st);
int etd_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt, double* coef,
               double f2w, hipStream_t st);

int etd_exp_run(int B, int n, float* u, float* v, long long pitch, long long istride, int ghosts, const float* coef,
                const InstParams* ip, int order, int prev, const float* nu, const float* nv, int n_steps,
                cpx<float>* work, hipStream_t st);
int etd_exp_run(int B, int n, double* u, double* v, long long pitch, long long istride, int ghosts,
                const double* coef, const InstParams* ip, int order, int prev, const double* nu, const double* nv,
                int n_steps, cpx<double>* work, hipStream_t st);
int etd_phi_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt,
                   float* coef, hipStream_t st);
int etd_phi_coeffs(int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv, double dt,
                   double* coef, hipStream_t st);

// phi_k(z) of the reference's jitted kernels (moljax/core/jit_kernels.py:28-53):
// the direct quotients, and the cubic Taylor polynomials below |z| = 1e-4.
__device__ __forceinline__ void phi123(double z, double& ez, double& p1, double& p2, double& p3) {
  ez = exp(z);
  if (fabs(z) < 1e-4) {
    p1 = 1.0 + z / 2.0 + z * z / 6.0 + z * z * z / 24.0;
    p2 = 0.5 + z / 6.0 + z * z / 24.0 + z * z * z / 120.0;
    p3 = 1.0 / 6.0 + z / 24.0 + z * z / 120.0 + z * z * z / 720.0;
  } else {
    p1 = (ez - 1.0) / z;
    p2 = (ez - 1.0 - z) / (z * z);
    p3 = (ez - 1.0 - z - z * z / 2.0) / (z * z * z);
  }
}

// Spectral multiplier tables of the jitted ETD kernels, elementwise in FP64
// from eigenvalues lam on an (ny, nx) index grid (ny = 1: one line) and step
// dt (z = dt lam):
//   mode 0  phi_1, phi_2, phi_3 of lam itself (z = lam)  -> o0..o2
//   mode 1  ETD1:   e^z, dt phi1                          -> o0, o1
//   mode 2  ETD2:   e^z, dt phi1 + dt phi2, -dt phi2      -> o0..o2
//   mode 3  ETDRK4: e^z, dt b1, dt b2 (= dt b3), dt b4     -> o0..o3
//           b1 = phi1 - 3 phi2 + 4 phi3, b2 = 2 phi2 - 4 phi3, b4 = -phi2 + 4 phi3
// With even != 0 every output is its even part (m(k) + m(-k)) / 2, indices
// reversed modulo ny and nx: the multiplier that the real part of
// ifft(m fft(x)) applies to a real x.
__device__ __forceinline__ void phi_outputs(double z, double dt, int mode, double* r) {
  double ez, p1, p2, p3;
  phi123(z, ez, p1, p2, p3);
  if (mode == 0) {
    r[0] = p1;
    r[1] = p2;
    r[2] = p3;
  } else if (mode == 1) {
    r[0] = ez;
    r[1] = dt * p1;
  } else if (mode == 2) {
    r[0] = ez;
    r[1] = dt * p1 + dt * p2;
    r[2] = -(dt * p2);
  } else {
    r[0] = ez;
    r[1] = dt * ((p1 - 3.0 * p2) + 4.0 * p3);
    r[2] = dt * (2.0 * p2 - 4.0 * p3);
    r[3] = dt * (-p2 + 4.0 * p3);
  }
}

__global__ __launch_bounds__(256) void k_phi_tables(int ny, int nx, const double* __restrict__ lam, double dt,
                                                    int mode, int even, double* __restrict__ o0,
                                                    double* __restrict__ o1, double* __restrict__ o2,
                                                    double* __restrict__ o3) {
  const long long n = (long long)ny * nx;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int ky = (int)(g / nx), kx = (int)(g - (long long)ky * nx);
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  phi_outputs(mode == 0 ? lam[g] : dt * lam[g], dt, mode, r);
  if (even) {
    const long long gm = (long long)((ny - ky) % ny) * nx + (nx - kx) % nx;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    phi_outputs(mode == 0 ? lam[gm] : dt * lam[gm], dt, mode, q);
    for (int i = 0; i < 4; ++i) r[i] = 0.5 * (r[i] + q[i]);
  }
  o0[g] = r[0];
  o1[g] = r[1];
  if (o2) o2[g] = r[2];
  if (o3) o3[g] = r[3];
}

static bool etd_size(int n) { return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024; }
// Square power-of-two grids run the Stockham ETD kernels; other grids with
// sides in [2, 2048] the chirp-z ones.
static bool etd_fast(int ny, int nx) { return ny == nx && etd_size(nx); }
static int etd_size_check(int ny, int nx, const char* what) {
  if (etd_fast(ny, nx) || (czt_len_ok(ny) && czt_len_ok(nx))) return MJX_OK;
  return fail(MJX_E_SIZE, std::string(what) + " grid " + std::to_string(ny) + "x" + std::to_string(nx) +
                              " not supported: sides must be in [2, 2048]");
}
}  // namespace mjx

using namespace mjx;

extern "C" {

long long mjx_etdrk4_coeff_elems(int ny, int nx) { return 6LL * ny * (nx + 2); }

long long mjx_etdrk4_workspace_bytes(int dtype, int B, int ny, int nx) {
  const long long es = dtype == MJX_F64 ? 16 : 8;
  if (!etd_fast(ny, nx)) return es * czt_etd_work_elems(B, ny, nx);
  // four packed planes (etd_planes in etd_kernels.h), the last one with room
  // for the separated N^ plane of ETD2
  return es * (long long)B * (4LL * ny * nx + 2LL * ny);
}

int mjx_etdrk4_coeffs_form(int dtype, int form, int ny, int nx, const double* lamx, const double* lamy, double Du,
                           double Dv, double dt, void* coef, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (form != 0 && form != 1) return fail(MJX_E_ARG, "form must be 0 (Cox-Matthews) or 1 (benchmark as written)");
  if (ny < 2 || nx < 2) return fail(MJX_E_ARG, "grid must be at least 2x2");
  if (!lamx || !lamy || !coef) return fail(MJX_E_ARG, "null pointer argument");
  hipStream_t st = (hipStream_t)stream;
  const double f2w = form == 0 ? 2.0 : 1.0;
  if (dtype == MJX_F32) return etd_coeffs(ny, nx, lamx, lamy, Du, Dv, dt, (float*)coef, f2w, st);
  return etd_coeffs(ny, nx, lamx, lamy, Du, Dv, dt, (double*)coef, f2w, st);
}

int mjx_etdrk4_coeffs(int dtype, int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv,
                      double dt, void* coef, void* stream) {
  return mjx_etdrk4_coeffs_form(dtype, 0, ny, nx, lamx, lamy, Du, Dv, dt, coef, stream);
}

int mjx_gs_etdrk4(int dtype, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                  int fill_ghosts, const void* coef, const double* inst, int n_steps, void* work, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (int rc = etd_size_check(ny, nx, "ETDRK4")) return rc;
  if (B < 1 || !u || !v || !coef || !inst || !work) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (pitch < nx || (fill_ghosts && pitch < nx + 2)) return fail(MJX_E_ARG, "pitch too small");
  if (n_steps == 0) return MJX_OK;
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (!etd_fast(ny, nx)) {
    if (dtype == MJX_F32)
      return czt_etdrk4_run(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)coef, ip,
                            n_steps, (cpx<float>*)work, st);
    return czt_etdrk4_run(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)coef, ip,
                          n_steps, (cpx<double>*)work, st);
  }
  if (dtype == MJX_F32)
    return etd_run(B, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)coef, ip, n_steps,
                   (cpx<float>*)work, st);
  return etd_run(B, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)coef, ip, n_steps,
                 (cpx<double>*)work, st);
}

int mjx_phi_tables(int ny, int nx, const double* lam, double dt, int mode, int even, double* out0, double* out1,
                   double* out2, double* out3, void* stream) {
  if (ny < 1 || nx < 1 || mode < 0 || mode > 3) return fail(MJX_E_ARG, "ny, nx must be >= 1 and mode in [0, 3]");
  if (!lam || !out0 || !out1 || (mode != 1 && !out2) || (mode == 3 && !out3))
    return fail(MJX_E_ARG, "null pointer argument");
  const long long n = (long long)ny * nx;
  hipLaunchKernelGGL(k_phi_tables, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, ny, nx, lam,
                     dt, mode, even, out0, out1, mode == 1 ? nullptr : out2, mode == 3 ? out3 : nullptr);
  return check_launch("phi_tables");
}

long long mjx_etd_phi_coeff_elems(int ny, int nx) { return 3LL * ny * (nx + 2); }

int mjx_etd_phi_coeffs(int dtype, int ny, int nx, const double* lamx, const double* lamy, double Du, double Dv,
                       double dt, void* coef, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (ny < 2 || nx < 2) return fail(MJX_E_ARG, "grid must be at least 2x2");
  if (!lamx || !lamy || !coef) return fail(MJX_E_ARG, "null pointer argument");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MJX_F32) return etd_phi_coeffs(ny, nx, lamx, lamy, Du, Dv, dt, (float*)coef, st);
  return etd_phi_coeffs(ny, nx, lamx, lamy, Du, Dv, dt, (double*)coef, st);
}

int mjx_gs_etd_exp(int dtype, int order, int B, int ny, int nx, void* u, void* v, long long pitch, long long istride,
                   int fill_ghosts, const void* coef, const double* inst, int prev, const void* nu_prev,
                   const void* nv_prev, int n_steps, void* work, void* stream) {
  if (dtype != MJX_F32 && dtype != MJX_F64) return fail(MJX_E_ARG, "dtype must be MJX_F32 or MJX_F64");
  if (order != 1 && order != 2) return fail(MJX_E_ARG, "order must be 1 (ETD1) or 2 (ETD2)");
  if (prev < 0 || prev > 2 || (order == 1 && prev != 0)) return fail(MJX_E_ARG, "prev must be 0, 1 or 2 (ETD2 only)");
  if (int rc = etd_size_check(ny, nx, "ETD")) return rc;
  if (B < 1 || !u || !v || !coef || !inst || !work) return fail(MJX_E_ARG, "null pointer or empty batch");
  if (prev == 1 && (!nu_prev || !nv_prev)) return fail(MJX_E_ARG, "prev = 1 needs nu_prev and nv_prev");
  if (n_steps < 0) return fail(MJX_E_ARG, "n_steps must be >= 0");
  if (pitch < nx || (fill_ghosts && pitch < nx + 2)) return fail(MJX_E_ARG, "pitch too small");
  if (n_steps == 0) return MJX_OK;
  int rc = ensure_tables(nullptr);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const InstParams* ip = reinterpret_cast<const InstParams*>(inst);
  if (!etd_fast(ny, nx)) {
    if (dtype == MJX_F32)
      return czt_etd_exp_run(B, ny, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)coef, ip,
                             order, prev, (const float*)nu_prev, (const float*)nv_prev, n_steps, (cpx<float>*)work,
                             st);
    return czt_etd_exp_run(B, ny, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)coef, ip,
                           order, prev, (const double*)nu_prev, (const double*)nv_prev, n_steps, (cpx<double>*)work,
                           st);
  }
  if (dtype == MJX_F32)
    return etd_exp_run(B, nx, (float*)u, (float*)v, pitch, istride, fill_ghosts, (const float*)coef, ip, order, prev,
                       (const float*)nu_prev, (const float*)nv_prev, n_steps, (cpx<float>*)work, st);
  return etd_exp_run(B, nx, (double*)u, (double*)v, pitch, istride, fill_ghosts, (const double*)coef, ip, order, prev,
                     (const double*)nu_prev, (const double*)nv_prev, n_steps, (cpx<double>*)work, st);
}

}  // extern "C"
