// mpres_dot_noalloc: mpres_dot with its partial results in a fixed device array, for
// callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The partials live in `scratch`, a file-scope device array of MPRES_NOALLOC_SCRATCH_BYTES.
// In host code the array's name is only a stand-in for the device symbol, so its device
// address is looked up once per device with hipGetSymbolAddress (at context creation,
// with the context's device current) and only that address is handed to kernels. Every
// GPU holds its own copy of the array until the process ends; all calls on one GPU share
// it, so calls that overlap in time on one GPU (different streams, or host threads) race.
// Nothing in the array has to start at zero: the dot's first pass writes every partial
// before any pass reads it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mpres_hip.h"
#include "internal.hpp"

namespace {

static __device__ __align__(256) char scratch[MPRES_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device, once looked up

}  // namespace

// Looks up the device address of `scratch` on `device`, which the caller has made current.
int mph::noalloc_init(int device) {
    if (device < 0 || device >= MAX_DEVICES)
        return set_error(MPRES_ERR_NO_DEVICE, "device %d beyond the %d devices the no-allocation path tracks", device,
                         MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return MPRES_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(MPRES_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_ptr[device] = p;
    return MPRES_OK;
}

extern "C" int mpres_dot_noalloc(mpres_ctx_t *c, int n, const mp_array_t *x, int incx, const mp_array_t *y, int incy,
                                 mp_array_t *r, int g3, int b3, void *st) {
    if (!c || !x || !y || !r) return mph::set_error(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;
    if (g3 <= 0 || b3 <= 0) return mph::set_error(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g3, b3);
    const size_t need = mpres_dot_workspace(c, n, g3, b3);
    if (need > (size_t)MPRES_NOALLOC_SCRATCH_BYTES)
        return mph::set_error(MPRES_ERR_WORKSPACE,
                              "reduction shape %d x %d needs %zu bytes of partials; the no-allocation dot holds at "
                              "most %d bytes",
                              g3, b3, need, MPRES_NOALLOC_SCRATCH_BYTES);
    const int dev = mph::ctx_device(c);
    void *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        p = (dev >= 0 && dev < MAX_DEVICES) ? g_ptr[dev] : nullptr;
    }
    if (!p) return mph::set_error(MPRES_ERR_HIP, "scratch address of device %d was not looked up", dev);
    // the caller-workspace path of mpres_dot: no pool, no allocation
    return mpres_dot(c, n, x, incx, y, incy, r, g3, b3, p, (size_t)MPRES_NOALLOC_SCRATCH_BYTES, st);
}
