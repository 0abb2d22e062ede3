// Kernel instantiation: 16 moduli, 16 lanes per number.
#define MPH_KERNELS 1
#include "launch.hpp"

MPH_INSTANTIATE(16, 16)
