// Kernel instantiations for 8 moduli (lanes per number: 2 8).
#define MPH_KERNELS 1
#include "launch.hpp"

namespace mph {

const Ops *ops_n8(int L) {
    static const Ops t2 = Launch<8, 2>::table();
    static const Ops t8 = Launch<8, 8>::table();
    switch (L) {
        case 2: return &t2;
        case 8: return &t8;
        default: return nullptr;
    }
}

}  // namespace mph
