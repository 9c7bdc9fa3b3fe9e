// Shared host/device helpers of the CPU-free solver library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/cfm.h"

namespace cfm {

constexpr int kWave = 64;  // CDNA wavefront width

// Status of the last failing HIP call is printed (stderr) and mapped to a
// cfm_status; nothing here aborts the process.
#define CFM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "cfm: HIP call \"%s\" failed at %s:%d: %s\n", #call,   \
                         __FILE__, __LINE__, hipGetErrorString(e_));                   \
            return (e_ == hipErrorOutOfMemory) ? CFM_ENOMEM : CFM_EDEVICE;             \
        }                                                                              \
    } while (0)

// ---------------------------------------------------------------- barrier --
// Device-wide barrier state for one cooperative launch.  `count` grows by
// one per workgroup per barrier (never reset inside a launch); barrier number
// e completes when count reaches e * gridDim.x.  `abort` is raised when a
// workgroup spins longer than the limit; every workgroup then leaves its loop
// at its next barrier, so the grid always drains.
struct GridSync {
    unsigned long long count;
    unsigned int abort;
    unsigned int pad0;
    unsigned long long spin_limit_ticks;  // s_memrealtime ticks (100 MHz)
    unsigned long long pad1[5];
};

__global__ void reset_sync_kernel(GridSync *s, unsigned long long limit_ticks);

// Returns true when the barrier completed, false when the launch is aborting.
__device__ __forceinline__ bool grid_barrier(GridSync *s, unsigned long long &epoch) {
    __shared__ int s_abort;
    __syncthreads();
    epoch += 1;
    if (threadIdx.x == 0) {
        int ab = 0;
        // Release this workgroup's stores to agent scope (L2 write-back
        // across XCDs), then arrive.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&s->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long target = epoch * (unsigned long long)gridDim.x;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        const unsigned long long limit = s->spin_limit_ticks;
        while (__hip_atomic_load(&s->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ab = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            if (__builtin_amdgcn_s_memrealtime() - t0 > limit) {
                __hip_atomic_store(&s->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_abort = ab;
    }
    __syncthreads();
    return s_abort == 0;
}

// --------------------------------------------------------- reductions ----
// xor butterfly over the 64 lanes of a wavefront: offsets 32,16,8,4,2,1.
// Every lane ends with the same value because a+b == b+a exactly.
__device__ __forceinline__ double wave_butterfly(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a 256-thread workgroup: butterfly per wavefront, then
// (w0 + w1) + (w2 + w3).  Result returned to every thread.
__device__ __forceinline__ double block256_sum(double v, double *lds4) {
    v = wave_butterfly(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();  // lds4 may still be read from a previous call
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    return (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
}

// ------------------------------------------------------------- launch ----
struct DeviceProps {
    int device = -1;
    int cus = 0;
    int coop = 0;
    char name[256] = {0};
};

cfm_status current_device_props(DeviceProps *out);

inline unsigned long long spin_ticks(int spin_limit_us) {
    const unsigned long long us = spin_limit_us > 0 ? (unsigned long long)spin_limit_us : 10000000ull;
    return us * 100ull;  // s_memrealtime runs at 100 MHz
}

}  // namespace cfm
