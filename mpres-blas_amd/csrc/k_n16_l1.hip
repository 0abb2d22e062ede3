// Kernel instantiation: 16 moduli, 1 lane per number.
#define MPH_KERNELS 1
#include "launch.hpp"

MPH_INSTANTIATE(16, 1)
