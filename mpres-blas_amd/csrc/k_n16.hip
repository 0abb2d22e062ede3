// Kernel instantiations for 16 moduli (lanes per number: 4 16).
#define MPH_KERNELS 1
#include "launch.hpp"

namespace mph {

const Ops *ops_n16(int L) {
    static const Ops t4 = Launch<16, 4>::table();
    static const Ops t16 = Launch<16, 16>::table();
    switch (L) {
        case 4: return &t4;
        case 16: return &t16;
        default: return nullptr;
    }
}

}  // namespace mph
