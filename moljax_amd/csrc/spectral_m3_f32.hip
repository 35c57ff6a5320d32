// FP32 periodic spectral launchers of the mixed-radix lengths 96, 192, 384, 768.
#include "spectral_kernels.h"

namespace mjx {

template struct Launch<float, SET_MIX3>;

}  // namespace mjx
