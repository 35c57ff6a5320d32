// Gray-Scott IMEX step chains on the periodic spectral kernels, for one
// batch of instances, from physical (u, v) in to physical (u, v) out.
//
// Every chain keeps the state in packed x-spectral form between kernels, so a
// step costs two HBM round trips of the batch (row pass + column pass):
//   STRANG  fwd_pack -> col(H) -> [row_gs(Heun) -> col(H^2)]*(n-1)
//           -> row_gs(Heun) -> col(H) -> inv_unpack,   H = (1 - dt/2 D lam)^-1
//           (moljax/core/stepping.py:370-435; adjacent diffusion half-steps
//           of consecutive steps are merged into one H^2 column pass)
//   EULER   fwd_pack -> [row_gs(Euler) -> col(H)]*n -> inv_unpack,
//           H = (1 - dt D lam)^-1   (stepping.py:317-367)
//   SSPRK2  per step: fwd_pack -> row(Euler) -> col(H) -> row(Euler) -> col(H)
//           -> inv_unpack(out = y/2 + E(E(y))/2), in place only
//           (stepping.py:438-495)
//   BENCH_STRANG  the Gray-Scott benchmark's splitting
//           (benchmarks/benchmark_gray_scott.py:268-289): reaction Euler
//           half-step, exact diffusion e^{dt D lam}, reaction Euler half-step:
//           fwd_pack -> row(half) -> col(E) -> [row(half, half) -> col(E)]*(n-1)
//           -> inv_unpack(closing half-step)
//   STRANG_INC  the STRANG operator in increment form, in place: the state
//           stays physical, held as its deviation (u - 1, v) from the steady
//           state during the chain, and every diffusion solve is
//           y + (H^k - 1) y, with only the increment going through the
//           transforms:
//           row_inc(first) -> col(H - 1) -> [row_inc(update, Heun,
//           transform) -> col(H^2 - 1)]*(n-1) -> row_inc(update, Heun,
//           transform) -> col(H - 1) -> row_inc(update)
//           (power-of-two sides; 48 B per point and step in FP32 against 32
//           for STRANG, with transform rounding that scales with the
//           increment instead of the state; run by gs_imex_inc_t in
//           spectral.hip, whose hot launches take the per-kernel events)
// The instance table's beta columns carry dt/2 D (STRANG, STRANG_INC) or
// dt D (EULER, SSPRK2, BENCH_STRANG); its dt column is the step size.
#pragma once
#include "mjx_common.h"
#include "spectral_api.h"

namespace mjx {

enum { GS_STRANG = 0, GS_EULER = 1, GS_SSPRK2 = 2, GS_BENCH_STRANG = 3, GS_STRANG_INC = 4 };

template <typename T>
int gs_chain_t(int scheme, int B, int ny, int nx, const T* u_in, const T* v_in, long long in_pitch,
               long long in_istride, T* u_out, T* v_out, long long out_pitch, long long out_istride, int ghosts,
               const T* lamx, const T* lamy, const T* lam2d, const InstParams* inst, int n_steps, cpx<T>* W,
               hipStream_t st) {
  int rc = MJX_OK;
#define MJX_TRY(x) \
  if ((rc = (x)) != MJX_OK) return rc
  if (scheme == GS_STRANG) {
    MJX_TRY(spec_row_fwd_pack(nx, B, ny, u_in, v_in, in_pitch, in_istride, W, st));
    MJX_TRY(spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 1, st));
    for (int s = 0; s < n_steps; ++s) {
      MJX_TRY(spec_row_gs(nx, B, ny, W, inst, st, RXN_HEUN));
      MJX_TRY(spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, s == n_steps - 1 ? 1 : 2, st));
    }
    return spec_row_inv_unpack(nx, B, ny, W, u_out, v_out, out_pitch, out_istride, ghosts, st);
  }
  // The column pass folds the whole 1/(nx ny) normalisation in, so a row
  // pass that directly follows fwd_pack gets its 1/nx share from the pack
  // (a power of two: exact).
  const double first_row_scale = 1.0 / (double)nx;
  if (scheme == GS_EULER) {
    MJX_TRY(spec_row_fwd_pack(nx, B, ny, u_in, v_in, in_pitch, in_istride, W, st, first_row_scale));
    for (int s = 0; s < n_steps; ++s) {
      MJX_TRY(spec_row_gs(nx, B, ny, W, inst, st, RXN_EULER));
      MJX_TRY(spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 1, st));
    }
    return spec_row_inv_unpack(nx, B, ny, W, u_out, v_out, out_pitch, out_istride, ghosts, st);
  }
  if (scheme == GS_BENCH_STRANG) {
    MJX_TRY(spec_row_fwd_pack(nx, B, ny, u_in, v_in, in_pitch, in_istride, W, st, first_row_scale));
    for (int s = 0; s < n_steps; ++s) {
      MJX_TRY(spec_row_gs(nx, B, ny, W, inst, st, s == 0 ? RXN_HALF : RXN_HALF2));
      MJX_TRY(spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, -2, st));
    }
    return spec_row_inv_unpack(nx, B, ny, W, u_out, v_out, out_pitch, out_istride, ghosts, st, 0.0, inst);
  }
  if (scheme == GS_SSPRK2) {
    if (u_in != u_out || v_in != v_out || in_pitch != out_pitch || in_istride != out_istride)
      return fail(MJX_E_ARG, "IMEX-SSPRK2 runs in place");
    for (int s = 0; s < n_steps; ++s) {
      MJX_TRY(spec_row_fwd_pack(nx, B, ny, u_out, v_out, out_pitch, out_istride, W, st, first_row_scale));
      for (int e = 0; e < 2; ++e) {
        MJX_TRY(spec_row_gs(nx, B, ny, W, inst, st, RXN_EULER));
        MJX_TRY(spec_col_apply(ny, B, nx, W, lamx, lamy, lam2d, inst, 1, st));
      }
      MJX_TRY(spec_row_inv_unpack(nx, B, ny, W, u_out, v_out, out_pitch, out_istride,
                                  s == n_steps - 1 ? ghosts : 0, st, 0.5));
    }
    return MJX_OK;
  }
#undef MJX_TRY
  return fail(MJX_E_ARG, "unknown IMEX scheme");
}

}  // namespace mjx
