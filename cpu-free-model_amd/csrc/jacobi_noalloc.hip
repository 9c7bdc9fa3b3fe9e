// Jacobi sweeps on caller-owned device fields with no allocation at all: the
// device-wide barrier state and the per-workgroup update-norm partials live in
// a file-scope device array (`scratch`) instead of plan buffers, so the call
// makes no hipMalloc / hipFree and creates no stream or event.  Launch
// configuration and kernels are the plan path's (jacobi_configure,
// jacobi_enqueue), so results are the same bits.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <mutex>

#include "jacobi_host.hpp"

namespace cfm {
namespace {

// One copy per GPU, held until the process ends: GridSync first, then the
// norm partials [slot][workgroup], then one FP64 sum per slot.
constexpr size_t kScratchBytes = size_t(2) << 20;
constexpr size_t kSyncBytes = (sizeof(GridSync) + 255) / 256 * 256;
constexpr size_t kScratchDoubles = (kScratchBytes - kSyncBytes) / sizeof(double);
static_assert(kSyncBytes < kScratchBytes, "scratch too small for the barrier state");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device; the host-side name of the array is not a device address).
cfm_status scratch_ptr(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return CFM_FAIL(CFM_EDEVICE, "jacobi_noalloc: device %d", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return CFM_FAIL(CFM_EDEVICE, "jacobi_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                            hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return CFM_OK;
}

// sums[s] = part[s][0] + part[s][1] + ... in workgroup order, the order the
// plan path adds them on the host (so the square roots are the same bits).
__global__ void norm_slot_sums_kernel(const double *part, int slots, int grid, double *sums) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    double acc = 0.0;
    for (int b = 0; b < grid; ++b) acc += part[(size_t)s * grid + b];
    sums[s] = acc;
}

}  // namespace
}  // namespace cfm

using namespace cfm;

extern "C" cfm_status cfm_jacobi_run_noalloc(const cfm_jacobi_params *p, void *field, void *work,
                                             void *stream_ptr, double *norms, int norm_cap, int *norm_count,
                                             cfm_timing *t) {
    if (!p || !field || !work || field == work) return CFM_FAIL(CFM_EINVAL, "jacobi_noalloc: two distinct fields");
    if (p->phase_timers) return CFM_FAIL(CFM_EINVAL, "jacobi_noalloc: phase timers need a plan");
    const auto w0 = std::chrono::steady_clock::now();
    JacobiConfig c;
    cfm_status st = jacobi_configure(p, &c);
    if (st != CFM_OK) return st;
    const size_t slots = p->variant == CFM_JACOBI_NO_COMPUTE ? 0 : (size_t)c.norm_slots;
    if (slots * (size_t)c.grid + slots > kScratchDoubles)
        return CFM_FAIL(CFM_EINVAL,
                        "jacobi_noalloc: %zu norm slots x %d workgroups exceed the %zu-double scratch array "
                        "(fewer norms, or a plan)", slots, c.grid, kScratchDoubles);
    unsigned char *base = nullptr;
    st = scratch_ptr(c.dev.device, &base);
    if (st != CFM_OK) return st;
    auto *sync = reinterpret_cast<GridSync *>(base);
    auto *part = reinterpret_cast<double *>(base + kSyncBytes);
    double *sums = part + slots * (size_t)c.grid;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    const size_t bytes = c.elems * c.elem_size;
    // the second working array starts as a copy of the field (faces and halos in both)
    CFM_HIP(hipMemcpyAsync(work, field, bytes, hipMemcpyDeviceToDevice, stream));
    int coop = -1;
    const auto k0 = std::chrono::steady_clock::now();
    st = jacobi_enqueue(*p, c, field, work, sync, part, nullptr, stream, &coop);
    if (st != CFM_OK) return st;
    CFM_HIP(hipStreamSynchronize(stream));
    const double kms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - k0).count();
    unsigned int aborted = 0;
    CFM_HIP(hipMemcpy(&aborted, &sync->abort, sizeof(aborted), hipMemcpyDeviceToHost));
    if (aborted) return barrier_timeout("jacobi_noalloc", coop);
    int produced = 0;
    if (slots > 0) {
        const int n = (int)slots;
        const int copy = norms ? std::min(n, std::max(norm_cap, 0)) : 0;
        if (copy > 0) {
            hipLaunchKernelGGL(norm_slot_sums_kernel, dim3((copy + 255) / 256), dim3(256), 0, stream, part, copy,
                               c.grid, sums);
            CFM_HIP(hipGetLastError());
            CFM_HIP(hipMemcpyAsync(norms, sums, sizeof(double) * copy, hipMemcpyDeviceToHost, stream));
            CFM_HIP(hipStreamSynchronize(stream));
            for (int s = 0; s < copy; ++s) norms[s] = std::sqrt(norms[s]);
        }
        produced = n;
    }
    if (norm_count) *norm_count = produced;
    // the result is in `work` after an odd number of passes: copy it back
    if (jacobi_passes(*p, c) & 1) {
        CFM_HIP(hipMemcpyAsync(field, work, bytes, hipMemcpyDeviceToDevice, stream));
        CFM_HIP(hipStreamSynchronize(stream));
    }
    jacobi_fill_timing(t, *p, c, (float)kms,
                       std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count(), coop);
    return CFM_OK;
}
