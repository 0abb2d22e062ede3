// Kernel instantiation: 32 moduli, 4 lanes per number.
#define MPH_KERNELS 1
#include "launch.hpp"

MPH_INSTANTIATE(32, 4)
