// Kernel instantiations for 32 moduli (lanes per number: 8 32).
#define MPH_KERNELS 1
#include "launch.hpp"

namespace mph {

const Ops *ops_n32(int L) {
    static const Ops t8 = Launch<32, 8>::table();
    static const Ops t32 = Launch<32, 32>::table();
    switch (L) {
        case 8: return &t8;
        case 32: return &t32;
        default: return nullptr;
    }
}

}  // namespace mph
