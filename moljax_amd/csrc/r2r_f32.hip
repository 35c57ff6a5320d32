// float instantiation of the DST-I / DCT line kernels.
#include "r2r_kernels.h"

namespace mjx {

int r2r_run(int kind, int orient, int mode, int L, R2RArgs a, int B, hipStream_t st, float) {
  return r2r_dispatch<float>(kind, orient, mode, L, a, B, st);
}

}  // namespace mjx
