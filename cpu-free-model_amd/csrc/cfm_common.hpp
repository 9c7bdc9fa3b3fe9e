// Shared host/device helpers of the CPU-free solver library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/cfm.h"

namespace cfm {

constexpr int kWave = 64;  // CDNA wavefront width

// Records the message returned by cfm_last_error() (per host thread).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Returns `code` after recording a message for cfm_last_error().
#define CFM_FAIL(code, ...) (::cfm::set_error(__VA_ARGS__), (code))

// A failing HIP call records its message and maps to a cfm_status; nothing
// here aborts the process.
#define CFM_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::cfm::set_error("HIP call \"%s\" failed at %s:%d: %s", #call, __FILE__, __LINE__,  \
                           hipGetErrorString(e_));                                            \
            return (e_ == hipErrorOutOfMemory) ? CFM_ENOMEM : CFM_EDEVICE;                    \
        }                                                                                     \
    } while (0)

// ---------------------------------------------------------------- barrier --
// Device-wide barrier for one cooperative launch, built from plain stores and
// loads (no atomic read-modify-write: device-scope atomics on one address
// serialise across the 8 XCDs and cost far more than the barrier itself).
//   arrive: workgroup b stores its barrier count e into arrive[b] (release);
//   gather: the 256 threads of workgroup 0 wait until every arrive[] >= e;
//   release: workgroup 0 stores e into 8 replicated `go` words (one per XCD,
//            each on its own 128-byte line), every workgroup polls its copy.
// Every wait is bounded by spin_limit_ticks of the 100 MHz s_memrealtime
// clock; a workgroup that runs out raises `abort`, and every workgroup leaves
// its loop at its next barrier, so the grid always drains.
constexpr int kMaxGrid = 8192;
constexpr int kGoCopies = 8;

struct GridSync {
    unsigned int go[kGoCopies * 32];  // copy i at go[32 * i]
    unsigned int abort;
    unsigned int pad0[31];
    unsigned long long spin_limit_ticks;
    unsigned long long pad1[15];
    unsigned int arrive[kMaxGrid];
};

__global__ void reset_sync_kernel(GridSync *s, unsigned long long limit_ticks, int grid);

__device__ __forceinline__ bool spin_expired(unsigned long long t0, unsigned long long limit) {
    return __builtin_amdgcn_s_memrealtime() - t0 > limit;
}

// Returns true when the barrier completed, false when the launch is aborting.
__device__ __forceinline__ bool grid_barrier(GridSync *s, unsigned int &epoch) {
    __shared__ int s_abort;
    __syncthreads();
    epoch += 1;
    const unsigned int e = epoch;
    const unsigned long long limit = s->spin_limit_ticks;
    if (threadIdx.x == 0) {
        s_abort = 0;
        __hip_atomic_store(&s->arrive[blockIdx.x], e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) {
        // gather: every thread of workgroup 0 watches a strided subset
        __syncthreads();
        int ab = 0;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        for (int b = threadIdx.x; b < (int)gridDim.x && !ab; b += blockDim.x) {
            while (__hip_atomic_load(&s->arrive[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < e) {
                if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    spin_expired(t0, limit)) {
                    __hip_atomic_store(&s->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (ab) s_abort = 1;
        __syncthreads();
        if (threadIdx.x < kGoCopies && !s_abort)
            __hip_atomic_store(&s->go[32 * threadIdx.x], e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        unsigned int *go = &s->go[32 * (blockIdx.x % kGoCopies)];
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < e) {
            if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                spin_expired(t0, limit)) {
                __hip_atomic_store(&s->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return s_abort == 0;
}

// Launches a persistent kernel (one that waits at grid_barrier).  A
// cooperative launch guarantees that every workgroup is resident at once.
// ROCm grants cooperative launches to one process at a time: while another
// process holds that grant the launch fails with hipErrorIllegalState, and
// the kernel then goes out as an ordinary launch.  Every grid here is sized to
// the co-resident maximum of an otherwise idle device, so all workgroups still
// start together; if one never became resident (another process filling the
// CUs) the barrier watchdog ends the launch and it reports CFM_ETIMEOUT.
// *cooperative (may be null) receives 1 for a cooperative launch, 0 for the
// ordinary-launch fallback; callers report it in cfm_timing.cooperative.
inline hipError_t launch_persistent(const void *fn, int grid, int block, void **args, hipStream_t stream,
                                    int *cooperative = nullptr) {
    hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, 0, stream);
    int coop = 1;
    if (e == hipErrorIllegalState) {
        (void)hipGetLastError();
        e = hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, stream);
        coop = 0;
    }
    if (cooperative) *cooperative = coop;
    return e;
}

// Status of a persistent launch whose barrier watchdog fired: the message
// says whether co-residency was guaranteed (cooperative launch) or only
// expected (ordinary-launch fallback), so a timeout can be told apart from a
// workgroup that never became resident.
inline cfm_status barrier_timeout(const char *what, int cooperative) {
    set_error("%s: a device-wide barrier exceeded its spin limit (%s)", what,
              cooperative ? "cooperative launch"
                          : "ordinary launch: another process held the cooperative-launch grant, so "
                            "co-residency of the workgroups was not guaranteed");
    return CFM_ETIMEOUT;
}

// -------------------------------------------------------- phase timers ----
// Per-workgroup time per phase (CFM_PHASE_*), from the constant-rate
// s_memrealtime clock read by thread 0 at every phase boundary: mark(p)
// charges the time since the previous boundary to phase p.  The state lives
// in the kernel's LDS (kClockWords words), so a timed build needs no extra
// registers in its loops; with ON = false every member is empty, the LDS
// array is unused and the timer code is compiled out.  flush() stores the
// workgroup's CFM_NPHASES sums and its first-to-last span (ticks) at
// out[blockIdx.x * kTimerSlots ...].
constexpr int kTimerSlots = CFM_NPHASES + 1;
constexpr int kClockWords = CFM_NPHASES + 2;

template <bool ON>
struct PhaseClock {
    unsigned long long *s;  // LDS: [0] first reading, [1] last reading, [2 + phase] sums
    __device__ __forceinline__ void start() {
        if constexpr (ON) {
            if (threadIdx.x == 0) {
                s[0] = s[1] = __builtin_amdgcn_s_memrealtime();
#pragma unroll
                for (int i = 0; i < CFM_NPHASES; ++i) s[2 + i] = 0;
            }
        }
    }
    __device__ __forceinline__ void mark(int phase) {
        if constexpr (ON) {
            if (threadIdx.x == 0) {
                const unsigned long long now = __builtin_amdgcn_s_memrealtime();
                s[2 + phase] += now - s[1];
                s[1] = now;
            }
        }
    }
    __device__ __forceinline__ void flush(unsigned long long *out) {
        if constexpr (ON) {
            if (out && threadIdx.x == 0) {
                unsigned long long *o = out + (size_t)blockIdx.x * kTimerSlots;
#pragma unroll
                for (int i = 0; i < CFM_NPHASES; ++i) o[i] = s[2 + i];
                o[CFM_NPHASES] = s[1] - s[0];
            }
        }
    }
};

// Host: per-workgroup tick sums -> cfm_phase_times (ticks at rate_khz).
inline void summarize_phases(const unsigned long long *ticks, int workgroups, double rate_khz, cfm_phase_times *out) {
    *out = cfm_phase_times{};
    if (workgroups < 1 || rate_khz <= 0.0) return;
    for (int b = 0; b < workgroups; ++b) {
        const unsigned long long *o = ticks + (size_t)b * kTimerSlots;
        for (int i = 0; i < CFM_NPHASES; ++i) {
            const double ms = (double)o[i] / rate_khz;
            out->ms[i] += ms / workgroups;
            if (ms > out->max_ms[i]) out->max_ms[i] = ms;
        }
        out->total_ms += (double)o[CFM_NPHASES] / rate_khz / workgroups;
    }
    out->workgroups = workgroups;
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

// Sum over the WAVES wavefronts of a workgroup, wave order 0, 1, 2, ...
// (used for update norms, which carry a tolerance, not a fixed tree).
template <int WAVES>
__device__ __forceinline__ double block_sum_waves(double v, double *lds) {
    v = wave_butterfly(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s += lds[w];
    return s;
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
