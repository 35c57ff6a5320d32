// Host-side API of the any-length (chirp-z) path, defined in czt.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "mjx_common.h"
#include "mjx_fft.h"

namespace mjx {

// Lengths served by the chirp-z kernels (convolution length <= 4096).
constexpr int CZT_MIN = 2, CZT_MAX = 2048;
inline bool czt_len_ok(int n) { return n >= CZT_MIN && n <= CZT_MAX; }

// The periodic spectral passes run the chirp-z kernels unless both sides
// have a Stockham plan: powers of two in [16, 4096] and the mixed-radix
// lengths of MJX_FOR_MIXED.
inline bool spec_generic(int nx, int ny) { return !(supported_len(nx) && supported_len(ny)); }

// Complex columns per row of the periodic spectral work array: nx for the
// packed Stockham layout, 2 (nx/2 + 1) for the half-spectra of the chirp-z
// layout.
inline long long spec_work_cols(int nx, int ny) { return spec_generic(nx, ny) ? 2LL * (nx / 2 + 1) : nx; }

// Chirp-z tables of one plan: n points, period P, doubled offsets K2 = 2 ko
// (output) and N2 = 2 no (input), convolution length L (a power of two,
// >= max(16, 2n - 1)); each table is [c (n) | d (n) | FFT_L(e)/L (L)].
struct CztTables {
  int n = 0, L = 0;
  const cpx<float>* t32 = nullptr;
  const cpx<double>* t64 = nullptr;
};
int czt_tables(int n, int P, int K2, int N2, CztTables& out);

// Create (once per process and device) the tables that the periodic passes,
// or the DST/DCT transforms of one kind, need for side n, so that later
// launches -- for example under stream capture -- only look them up.
int czt_prepare_periodic(int n);
int czt_prepare_r2r(int kind, int n);

#define MJX_CZT_DECL(T)                                                                                       \
  int czt_row_fwd_pack(int nx, int B, int ny, const T* a, const T* b, long long pitch, long long istride,     \
                       cpx<T>* W, hipStream_t st, double in_scale);                                           \
  int czt_row_inv_unpack(int nx, int B, int ny, const cpx<T>* W, T* a, T* b, long long pitch,                \
                         long long istride, int ghosts, hipStream_t st, double keep, const InstParams* react); \
  int czt_row_gs(int nx, int B, int ny, cpx<T>* W, const InstParams* ip, hipStream_t st, int rxn, int mstep); \
  int czt_col_apply(int ny, int B, int nx, cpx<T>* W, const T* lamx, const T* lamy, const T* lam2d,          \
                    const InstParams* ip, int nrep, hipStream_t st, int mstep);                               \
  int czt_etdrk4_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,       \
                     const T* coef, const InstParams* ip, int n_steps, cpx<T>* work, hipStream_t st);          \
  int czt_etd_exp_run(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,      \
                      const T* coef, const InstParams* ip, int order, int prev, const T* nu, const T* nv,      \
                      int n_steps, cpx<T>* work, hipStream_t st);                                              \
  int czt_etdrk4_pack(int B, int ny, int nx, const T* u, const T* v, long long pitch, long long istride,      \
                      cpx<T>* work, hipStream_t st);                                                           \
  int czt_etdrk4_advance(int B, int ny, int nx, const T* coef, const InstParams* ip, int n_steps,             \
                         cpx<T>* work, hipStream_t st);                                                        \
  int czt_etdrk4_unpack(int B, int ny, int nx, T* u, T* v, long long pitch, long long istride, int ghosts,    \
                        cpx<T>* work, hipStream_t st);

MJX_CZT_DECL(float)
MJX_CZT_DECL(double)
#undef MJX_CZT_DECL

// Complex elements of the any-length ETD workspace (six planes of
// B * ny * (nx + 2)).
inline long long czt_etd_work_elems(int B, int ny, int nx) { return 6LL * B * ny * (nx + 2); }

// Non-periodic line work (kind 0 DST-I, 1 DCT-II, 2 DST-II), mode R2R_FWD,
// R2R_INV or R2R_APPLY, over B instances of nlines lines of n points.
struct CztR2RCall {
  int dtype, kind, mode, B, nlines, n;
  const void* in;
  long long in_es, in_ls, in_is;
  void* out;
  long long out_es, out_ls, out_is;
  double scale;             // FWD / INV output scale
  const void* lam_line;     // APPLY: symbol along the line
  const void* lam_cross;    // APPLY: symbol of the other dimension per line (or null)
  int n_cross;              // APPLY: the other dimension's length, 0 = none
  const double* inst;       // per-instance rows, or null (beta0, alpha 0)
  int mmode;
  double beta0;
  int accumulate;
  const double* lift;       // left, right, bottom, top contributions, or null
};
int czt_r2r(const CztR2RCall& c, hipStream_t st);

// Per-precision launch of the real line kernel (czt_f32.hip / czt_f64.hip).
struct CztR2RArgs;
int czt_r2r_launch(const CztR2RArgs& a, int orient, bool apply, const CztTables& h1, const CztTables& h2, int B,
                   hipStream_t st, float);
int czt_r2r_launch(const CztR2RArgs& a, int orient, bool apply, const CztTables& h1, const CztTables& h2, int B,
                   hipStream_t st, double);

}  // namespace mjx
