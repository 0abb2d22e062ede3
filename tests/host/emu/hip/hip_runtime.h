// Host stand-in for <hip/hip_runtime.h> used ONLY by the host tests
// (tests/host/kernels_emu.cpp, tests/host/batch_emu.cpp), which compile this
// library's own kernel templates with g++ and run them on the CPU, one lane per
// number (L = 1). With L = 1 a group is a single lane, so every cross-lane
// exchange is the identity. Kernels without workgroup synchronisation run their
// threads one after another; for kernels that synchronise (kernels_emu.cpp's
// run_wg, batch_emu.cpp) the threads of one workgroup run as concurrent host
// threads and emu_barrier is set to a barrier wait, so __syncthreads synchronises
// them.
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

struct int2 {
    int x, y;
};
struct int4 {
    int x, y, z, w;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline int __clz(int v) { return v ? __builtin_clz((unsigned)v) : 32; }
inline int __ffs(int v) { return __builtin_ffs(v); }

struct EmuIdx {
    unsigned x, y, z;
};
inline thread_local EmuIdx blockIdx, threadIdx, gridDim{1, 1, 1};
inline void (*emu_barrier)() = nullptr;

inline unsigned __lane_id() { return threadIdx.x & 63u; }
template <class T>
inline T __shfl_xor(T v, int, int = 64) { return v; }
template <class T>
inline T __shfl(T v, int, int = 64) { return v; }
template <class T>
inline T __shfl_down(T v, unsigned, int = 64) { return v; }
inline int atomicMin(int *p, int v) {
    int o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
inline int atomicMax(int *p, int v) {
    int o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline void __syncthreads() {
    if (emu_barrier) emu_barrier();
}
inline void __threadfence_block() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#define __builtin_amdgcn_readfirstlane(v) (v)
