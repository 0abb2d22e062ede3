// Host stand-in for <hip/hip_runtime.h> used ONLY by tests/host/kernels_emu.cpp,
// which compiles this library's own kernel templates with g++ and runs them
// sequentially on the CPU, one lane per number (L = 1). With L = 1 a group is
// a single lane, so every cross-lane exchange is the identity; kernels that
// synchronise a workgroup (LDS trees) are not emulated.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)
#define __shared__ static

typedef int hipError_t;
typedef void *hipStream_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };

struct int4 {
    int x, y, z, w;
};

struct EmuIdx {
    unsigned x, y, z;
};
extern EmuIdx blockIdx, threadIdx;

inline unsigned __lane_id() { return threadIdx.x & 63u; }
template <class T>
inline T __shfl_xor(T v, int, int = 64) { return v; }
template <class T>
inline T __shfl(T v, int, int = 64) { return v; }
inline int atomicMin(int *p, int v) {
    const int o = *p;
    if (v < o) *p = v;
    return o;
}
inline void __syncthreads() {}
inline void __threadfence_block() {}
