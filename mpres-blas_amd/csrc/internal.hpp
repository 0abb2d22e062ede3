// Internal declarations shared by the host-side translation units.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "devconsts.hpp"

struct mpres_ctx;

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

// error message of the calling thread (mpres_last_error); returns code
int set_error(int code, const char *fmt, ...);
// device of a context
int ctx_device(const struct ::mpres_ctx *c);
// looks up the no-allocation dot's scratch array on a device (current), once per device
int noalloc_init(int device);

}  // namespace mph
