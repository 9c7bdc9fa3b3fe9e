// Library-level entry points: version, status strings, device queries and
// the host-side generators (initial fields, Poisson CSR, seeded uniforms).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cfm_common.hpp"

extern "C" const char *cfm_version(void) { return "cfm 0.1.0 (gfx950, HIP)"; }

static thread_local char g_last_error[1024] = "";

void cfm::set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

extern "C" const char *cfm_last_error(void) { return g_last_error; }

extern "C" const char *cfm_status_str(int s) {
    switch (s) {
        case CFM_OK: return "ok";
        case CFM_EINVAL: return "invalid argument";
        case CFM_ENOMEM: return "out of memory";
        case CFM_EDEVICE: return "device error";
        case CFM_ETIMEOUT: return "device barrier timed out";
        case CFM_ENOTCONV: return "not converged";
        case CFM_EBREAKDOWN: return "CG breakdown (matrix not SPD)";
        case CFM_ENOTSUP: return "not supported";
        default: return "unknown status";
    }
}

extern "C" cfm_status cfm_device_count(int *count) {
    if (!count) return CFM_EINVAL;
    *count = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return CFM_EDEVICE;
    *count = n;
    return CFM_OK;
}

extern "C" cfm_status cfm_device_info(int dev, char *name, int name_cap, int *compute_units,
                                      size_t *mem_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || dev < 0 || dev >= n) return CFM_EDEVICE;
    hipDeviceProp_t prop;
    CFM_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_cap > 0) std::snprintf(name, name_cap, "%s (%s)", prop.name, prop.gcnArchName);
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (mem_bytes) *mem_bytes = prop.totalGlobalMem;
    return CFM_OK;
}

extern "C" cfm_status cfm_device_synchronize(void) {
    CFM_HIP(hipDeviceSynchronize());
    return CFM_OK;
}

// ---------------------------------------------------- bandwidth probe ---
// Plain streaming copy (16 B per lane per step, grid-stride).  It calibrates
// the hardware counters (FETCH_SIZE / WRITE_SIZE per known byte count) and
// gives the practical HBM ceiling next to which the solver kernels are read.
namespace {
typedef double v2f64_t __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) copy_probe_kernel(const v2f64_t *__restrict__ a,
                                                         v2f64_t *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// Shapes of the ceiling sweep: KIND 0 copy, 1 read only (a sum per workgroup, so the loads are
// not removed), 2 write only; U 16-byte vectors in flight per lane (grid-strided); NT
// non-temporal loads and stores.
template <int KIND, int U, bool NT>
__global__ void __launch_bounds__(256) shape_probe_kernel(const v2f64_t *__restrict__ a, v2f64_t *__restrict__ b,
                                                          size_t n, double *__restrict__ sums) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    v2f64_t acc = {0.0, 0.0};
    const v2f64_t fill = {1.0, 2.0};
    for (size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i0 < n; i0 += stride * U) {
        v2f64_t v[U];
        if constexpr (KIND != 2) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const size_t i = i0 + u * stride;
                if (i < n) v[u] = NT ? __builtin_nontemporal_load(a + i) : a[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + u * stride;
            if (i >= n) continue;
            if constexpr (KIND == 1) {
                acc += v[u];
            } else {
                const v2f64_t w = KIND == 0 ? v[u] : fill;
                if constexpr (NT)
                    __builtin_nontemporal_store(w, b + i);
                else
                    b[i] = w;
            }
        }
    }
    if constexpr (KIND == 1) {
        __shared__ double red[256];
        red[threadIdx.x] = acc[0] + acc[1];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
    }
}

using ShapeKernel = void (*)(const v2f64_t *, v2f64_t *, size_t, double *);
template <int KIND>
static ShapeKernel shape_kernel(int unroll, bool nt) {
    if (unroll == 1) return nt ? shape_probe_kernel<KIND, 1, true> : shape_probe_kernel<KIND, 1, false>;
    if (unroll == 4) return nt ? shape_probe_kernel<KIND, 4, true> : shape_probe_kernel<KIND, 4, false>;
    return nullptr;
}
}  // namespace

extern "C" cfm_status cfm_bandwidth_probe_shape(size_t bytes, int reps, int kind, int grid, int unroll, int nt,
                                                double *gbs, double *best_ms) {
    if (bytes < 16 || reps < 1 || !gbs || kind < 0 || kind > 2 || grid < 1 || grid > (1 << 20))
        return CFM_FAIL(CFM_EINVAL, "bandwidth probe: bytes >= 16, reps >= 1, kind 0..2, grid 1..2^20");
    ShapeKernel k = kind == 0 ? shape_kernel<0>(unroll, nt != 0)
                    : kind == 1 ? shape_kernel<1>(unroll, nt != 0)
                                : shape_kernel<2>(unroll, nt != 0);
    if (!k) return CFM_FAIL(CFM_EINVAL, "bandwidth probe: unroll 1 or 4");
    const size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    double *sums = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    cfm_status st = CFM_OK;
    float best = 1e30f;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
        hipMalloc((void **)&sums, sizeof(double) * grid) != hipSuccess) {
        st = CFM_ENOMEM;
    } else if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        st = CFM_EDEVICE;
    } else {
        for (int r = 0; r < reps + 1 && st == CFM_OK; ++r) {
            const hipError_t r0 = hipEventRecord(e0, nullptr);
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, nullptr, static_cast<const v2f64_t *>(a),
                               static_cast<v2f64_t *>(b), n, sums);
            const hipError_t r1 = hipEventRecord(e1, nullptr);
            float ms = 0.f;
            if (r0 != hipSuccess || r1 != hipSuccess || hipGetLastError() != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                st = CFM_EDEVICE;
            else if (r > 0 && ms < best)
                best = ms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sums) (void)hipFree(sums);
    if (st != CFM_OK) return st;
    *gbs = (kind == 0 ? 2.0 : 1.0) * (double)(n * 16) / (best * 1e-3) / 1e9;
    if (best_ms) *best_ms = best;
    return CFM_OK;
}

extern "C" cfm_status cfm_bandwidth_probe(size_t bytes, int reps, double *copy_gbs, double *best_ms) {
    if (bytes < 16 || reps < 1 || !copy_gbs) return CFM_EINVAL;
    const size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    cfm_status st = CFM_OK;
    float best = 1e30f;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        st = CFM_ENOMEM;
    } else if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        st = CFM_EDEVICE;
    } else {
        for (int r = 0; r < reps + 1 && st == CFM_OK; ++r) {
            const hipError_t r0 = hipEventRecord(e0, nullptr);
            hipLaunchKernelGGL(copy_probe_kernel, dim3(1024), dim3(256), 0, nullptr,
                               static_cast<const v2f64_t *>(a), static_cast<v2f64_t *>(b), n);
            const hipError_t r1 = hipEventRecord(e1, nullptr);
            float ms = 0.f;
            if (r0 != hipSuccess || r1 != hipSuccess || hipGetLastError() != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                st = CFM_EDEVICE;
            else if (r > 0 && ms < best)
                best = ms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (st != CFM_OK) return st;
    *copy_gbs = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
    if (best_ms) *best_ms = best;
    return CFM_OK;
}

// ------------------------------------------------------- initial fields ---
template <typename T>
static void init_field(int nx, int ny, int nz, int kind, T *a) {
    const size_t plane = (size_t)nx * ny;
    std::memset(a, 0, sizeof(T) * plane * (nz >= 3 ? nz : 1));
    // pi has the element type, as the reference's `const real PI{static_cast<real>(2.0 *
    // std::asin(1.0))}` (Stencil/jacobi2D/include/common.h:24): float-rounded for the
    // FP32 field, then widened to double in the sin argument
    const double pi = (double)(T)(2.0 * std::asin(1.0));
    if (nz < 3) {
        // Stencil/jacobi2D/src/common.cu:15-26 (offset 0)
        for (int iy = 0; iy < ny; ++iy) {
            const double y0 = std::sin(2.0 * pi * iy / (ny - 1));
            a[(size_t)iy * nx] = (T)y0;
            a[(size_t)iy * nx + nx - 1] = (T)y0;
        }
        return;
    }
    if (kind == CFM_INIT_REF_QUIRK) {
        // Stencil/jacobi3D/src/common.cu:18-28: the x=0 face keeps the value of
        // the last ix iteration, (iz - iy) - (nx - 1).
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                a[iz * plane + (size_t)iy * nx] = (T)(((double)iz - (double)iy) - (double)(nx - 1));
        return;
    }
    for (int iz = 0; iz < nz; ++iz) {
        const double sz = std::sin(2.0 * pi * iz / (nz - 1));
        for (int iy = 0; iy < ny; ++iy) {
            const double v = std::sin(2.0 * pi * iy / (ny - 1)) * sz;
            a[iz * plane + (size_t)iy * nx] = (T)v;
            a[iz * plane + (size_t)iy * nx + nx - 1] = (T)v;
        }
    }
}

extern "C" cfm_status cfm_jacobi_init_field(int nx, int ny, int nz, int kind, int prec, void *out) {
    if (!out || nx < 3 || ny < 3 || nz < 1 || nz == 2) return CFM_EINVAL;
    if (kind != CFM_INIT_SIN && kind != CFM_INIT_REF_QUIRK) return CFM_EINVAL;
    if (prec == CFM_F64)
        init_field<double>(nx, ny, nz, kind, static_cast<double *>(out));
    else if (prec == CFM_F32)
        init_field<float>(nx, ny, nz, kind, static_cast<float *>(out));
    else
        return CFM_EINVAL;
    return CFM_OK;
}

// ----------------------------------------------------------- generators ---
extern "C" int64_t cfm_gen_poisson3d(int nx, int ny, int nz, int *rowptr, int *col, double *val) {
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    const long long n = (long long)nx * ny * nz;
    if (n * 7 > 0x7fffffffLL) return -1;
    const long long plane = (long long)nx * ny;
    long long k = 0, row = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix, ++row) {
                if (rowptr) rowptr[row] = (int)k;
                // ascending column order: z-1, y-1, x-1, diagonal, x+1, y+1, z+1
                const long long cols[7] = {row - plane, row - nx, row - 1, row, row + 1, row + nx, row + plane};
                const bool ok[7] = {iz > 0, iy > 0, ix > 0, true, ix + 1 < nx, iy + 1 < ny, iz + 1 < nz};
                for (int j = 0; j < 7; ++j) {
                    if (!ok[j]) continue;
                    if (col) col[k] = (int)cols[j];
                    if (val) val[k] = j == 3 ? 6.0 : -1.0;
                    ++k;
                }
            }
    if (rowptr) rowptr[row] = (int)k;
    return k;
}

namespace {
// separate, non-inlined calls so that no compiler fuses sin and cos of one
// angle into sincos(), whose last bit can differ from the separate calls
__attribute__((noinline)) double host_cos(double a) { return std::cos(a); }
__attribute__((noinline)) double host_sin(double a) { return std::sin(a); }

struct SplitMix {
    uint64_t s;
    double next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    }
};
}  // namespace

// Draws of one variable-coefficient system: box lengths, then k per cell,
// then f per cell, all from the SplitMix64 stream (seed, system_id).
static bool varcoef_draw(uint64_t seed, int system_id, int lmin, int lmax, int L[3], std::vector<double> *k,
                         double *f) {
    if (lmin < 1 || lmax < lmin || system_id < 0) return false;
    SplitMix g{seed ^ (0x9E3779B97F4A7C15ULL * ((uint64_t)system_id + 1))};
    for (int a = 0; a < 3; ++a) {
        const double u = g.next();
        int v = (int)std::floor(std::exp(std::log((double)lmin) + u * (std::log((double)lmax + 1.0) - std::log((double)lmin))));
        L[a] = v < lmin ? lmin : v > lmax ? lmax : v;
    }
    const long long n = (long long)L[0] * L[1] * L[2];
    if (n * 7 > 0x7fffffffLL) return false;
    if (!k) return true;
    k->resize((size_t)n);
    for (long long i = 0; i < n; i += 2) {
        const double u1 = g.next(), u2 = g.next();
        const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        const double th = 2.0 * 3.14159265358979323846 * u2;
        (*k)[i] = std::exp(r * host_cos(th));
        if (i + 1 < n) (*k)[i + 1] = std::exp(r * host_sin(th));
    }
    for (long long i = 0; i < n; ++i) {
        const double u = g.next();
        if (f) f[i] = 2.0 * u - 1.0;
    }
    return true;
}

extern "C" cfm_status cfm_gen_varcoef3d_coef(uint64_t seed, int system_id, int lmin, int lmax, int *dims,
                                             double *kcell) {
    if (!dims || !kcell) return CFM_EINVAL;
    int L[3];
    std::vector<double> k;
    if (!varcoef_draw(seed, system_id, lmin, lmax, L, &k, nullptr)) return CFM_EINVAL;
    for (int a = 0; a < 3; ++a) dims[a] = L[a];
    std::memcpy(kcell, k.data(), sizeof(double) * k.size());
    return CFM_OK;
}

extern "C" int64_t cfm_gen_varcoef3d(uint64_t seed, int system_id, int lmin, int lmax, int *dims, int *rowptr,
                                     int *col, double *val, double *f) {
    if (!dims) return -1;
    int L[3];
    const bool sizes_only = !rowptr && !col && !val && !f;
    std::vector<double> k;
    if (!varcoef_draw(seed, system_id, lmin, lmax, L, sizes_only ? nullptr : &k, f)) return -1;
    for (int a = 0; a < 3; ++a) dims[a] = L[a];
    const long long nx = L[0], ny = L[1], nz = L[2], n = nx * ny * nz, plane = nx * ny;
    if (sizes_only) {
        long long nnz = 0;
        for (long long i = 0; i < n; ++i) {
            const long long ix = i % nx, iy = (i / nx) % ny, iz = i / plane;
            nnz += 1 + (iz > 0) + (iy > 0) + (ix > 0) + (ix + 1 < nx) + (iy + 1 < ny) + (iz + 1 < nz);
        }
        return nnz;
    }
    long long e = 0;
    for (long long i = 0; i < n; ++i) {
        const long long ix = i % nx, iy = (i / nx) % ny, iz = i / plane;
        const long long nb[6] = {iz > 0 ? i - plane : -1, iy > 0 ? i - nx : -1,         ix > 0 ? i - 1 : -1,
                                 ix + 1 < nx ? i + 1 : -1, iy + 1 < ny ? i + nx : -1, iz + 1 < nz ? i + plane : -1};
        double cf[6], diag = 0.0;
        for (int d = 0; d < 6; ++d) {
            cf[d] = nb[d] >= 0 ? (2.0 * k[i] * k[nb[d]]) / (k[i] + k[nb[d]]) : k[i];
            diag = diag + cf[d];
        }
        if (rowptr) rowptr[i] = (int)e;
        for (int d = 0; d < 7; ++d) {  // ascending columns; the diagonal sits between x-1 and x+1
            long long c;
            double v;
            if (d == 3) {
                c = i;
                v = diag;
            } else {
                const int dd = d < 3 ? d : d - 1;
                if (nb[dd] < 0) continue;
                c = nb[dd];
                v = -cf[dd];
            }
            if (col) col[e] = (int)c;
            if (val) val[e] = v;
            ++e;
        }
    }
    if (rowptr) rowptr[n] = (int)e;
    return e;
}

extern "C" cfm_status cfm_gen_uniform(uint64_t seed, uint64_t stream_id, double *out, int64_t count) {
    if (!out || count < 0) return CFM_EINVAL;
    // SplitMix64 (Steele, Lea & Flood, OOPSLA 2014); stream state
    // seed ^ (golden-ratio increment * (stream_id + 1)); 53-bit uniforms.
    uint64_t s = seed ^ (0x9E3779B97F4A7C15ULL * (stream_id + 1));
    for (int64_t i = 0; i < count; ++i) {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        out[i] = (double)(z >> 11) * 0x1.0p-53;
    }
    return CFM_OK;
}
