// Kernel instantiation: 32 moduli, 8 lanes per number.
#define MPH_KERNELS 1
#include "launch.hpp"

MPH_INSTANTIATE(32, 8)
