// Internal declarations shared by the host-side translation units.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "devconsts.hpp"

namespace mph {

struct ModuliSet {
    const char *name;
    int n;
    int log2m;
    int m[MAXN];
};

int num_sets();
const ModuliSet *get_set(int id);
bool compute_consts(const ModuliSet &s, DevConsts &c, std::vector<int> &pow2);

}  // namespace mph
