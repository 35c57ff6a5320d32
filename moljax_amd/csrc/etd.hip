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
  return es * (long long)B * (3LL * ny * nx + 3LL * ny * (nx + 2));
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
