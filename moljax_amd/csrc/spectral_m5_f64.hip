// FP64 periodic spectral launchers of the mixed-radix lengths 100, 200, 250, 500, 1000.
#include "spectral_kernels.h"

namespace mjx {

template struct Launch<double, SET_MIX5>;

}  // namespace mjx
