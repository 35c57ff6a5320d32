// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/moljax_hip.h"

namespace mjx {

// Per-instance parameters, laid out as 8 float64 values per instance
// (the host builds a (B, 8) float64 device array):
//   beta_a, beta_b   spectral denominator 1 - beta*lambda (field a / b)
//   alpha_a, alpha_b spectral numerator   1 + alpha*lambda (field a / b)
//   F, k, dt         Gray-Scott feed, kill and step
//   aux              unused (alignment)
struct InstParams {
  double beta_a, beta_b;
  double alpha_a, alpha_b;
  double F, k, dt, aux;
};

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
int ensure_tables(hipStream_t stream);
// Resolves the current device's address of the allocation-free entry point's
// scratch array (spectral.hip); called by mjx_init.
int scratch_setup();
// True when the current device's twiddle tables exist (mjx_init has run).
bool tables_ready();

// Tile shapes chosen by the autotuner: lines (rows or columns) per
// workgroup for a kernel family, precision and transform length; 0 means the
// built-in default.  Set by mjx_autotune_* or mjx_tune_set, read at launch.
enum { TK_ROW = 0, TK_COL = 1, TK_ETD_ROW = 2, TK_ETD_COL = 3, TK_NUM = 4 };
int tuned_lines(int kind, int dtype, int n);
void set_tuned_lines(int kind, int dtype, int n, int lines);

inline bool is_pow2_len(int n) {
  return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 ||
         n == 4096;
}

// Workgroup index remap so that the G workgroups of a launch are split into 8
// contiguous logical ranges, one per XCD (hardware dispatches workgroup w to
// XCD w mod 8).  Identity when G is not a multiple of 8.
__device__ __forceinline__ int xcd_remap(int w, int G) {
  if ((G & 7) != 0) return w;
  return (w & 7) * (G >> 3) + (w >> 3);
}

// The same tiles in the opposite order, for column passes: the row pass
// before them wrote the last instances last, and those are the ones still
// held by the Infinity Cache when the column pass starts.
__device__ __forceinline__ int xcd_remap_rev(int w, int G) { return G - 1 - xcd_remap(w, G); }

}  // namespace mjx
