// Persistent Jacobi sweeps (2D 5-point, 3D 7-point) for gfx950.
//
// Reference behaviour restated (arithmetic and loop semantics):
//   2D update  Stencil/jacobi2D/src/common.cu:36-41   new = 0.25*(E + W + N + S)
//   3D update  Stencil/jacobi3D/src/common.cu:41-47   new = (1/6)*(x+1 + x-1 + y+1 + y-1 + z+1 + z-1)
//   sweep loop, periodic halo copies and swap:
//              Stencil/jacobi2D/src/common.cu:214-253, Stencil/jacobi3D/src/common.cu:174-219
//   CPU-free scheme (one cooperative launch, device-wide barrier per sweep):
//              Stencil/jacobi3D/src/single-stream/multi-threaded-two-block-comm.cu:10-124
//
// Design (MI355X-first, not a translation of the CUDA blocks):
//   * A wavefront owns a tile of 64*VX consecutive x (VX = 16 B / sizeof(T),
//     one dwordx4 per lane per row) times RY rows of y, and streams the
//     slowest axis (z in 3D, y in 2D) through a chunk, keeping planes z-1, z,
//     z+1 in registers: every element is fetched from HBM once per sweep.
//   * x-neighbours come from the neighbouring lane (__shfl_up/__shfl_down);
//     only lanes 0 and 63 load one edge value per row.  y-neighbours come
//     from the wave's own rows plus two halo rows per plane (L2 hits).
//   * Workgroups are 4 independent wavefronts; the persistent grid holds
//     every co-resident workgroup, and the tile -> wavefront map keeps
//     neighbouring tiles on the same XCD (workgroup i runs on XCD i % 8).
//   * Compiled with -ffp-contract=off: the sums are evaluated left to right
//     exactly as the reference expression, so FP64 results are bit-exact
//     against the CPU oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cfm_common.hpp"

namespace cfm {

__global__ void reset_sync_kernel(GridSync *s, unsigned long long limit_ticks) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->count = 0;
        s->abort = 0;
        s->spin_limit_ticks = limit_ticks;
    }
}

struct JacobiArgs {
    void *buf0;
    void *buf1;
    int nx, ny, nz;
    int niter;        // sweeps in this launch
    int it0;          // index of the first sweep (parity and norm slots)
    int bc;
    int norm_every;
    int chunks;       // chunks along the streamed axis
    int chunk_len;
    int ntx, nty;     // tiles along x and y (nty = 1 in 2D)
    int no_compute;
    double *norm_part;  // [slot][gridDim.x]
    GridSync *sync;
};

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
    using type = double2;
};
template <>
struct Vec16<float> {
    using type = float4;
};

template <typename T>
__device__ __forceinline__ T stencil_coeff3d() {
    return T(1) / T(6);
}

// Load VX consecutive elements starting at p (x = first index), masked by nx.
template <typename T, int VX, bool ALIGNED>
__device__ __forceinline__ void load_row(const T *__restrict__ p, int x, int nx, bool row_ok,
                                         T (&v)[VX]) {
    if (ALIGNED) {
        if (row_ok && x + VX <= nx) {
            typename Vec16<T>::type t = *reinterpret_cast<const typename Vec16<T>::type *>(p + x);
            static_assert(sizeof(t) == VX * sizeof(T), "vector width");
            const T *tp = reinterpret_cast<const T *>(&t);
#pragma unroll
            for (int e = 0; e < VX; ++e) v[e] = tp[e];
            return;
        }
    }
#pragma unroll
    for (int e = 0; e < VX; ++e) v[e] = (row_ok && x + e < nx) ? p[x + e] : T(0);
}

// Maps the workgroup index so that consecutive logical workgroups share an
// XCD; returns the global wavefront index.
__device__ __forceinline__ int xcd_wave_id() {
    const int g = gridDim.x;
    const int b = blockIdx.x;
    int lb = b;
    if ((g & 7) == 0) lb = (b & 7) * (g >> 3) + (b >> 3);
    return lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
}

// ------------------------------------------------------------------ 3D ----
template <typename T, int RY, bool ALIGNED>
__device__ __forceinline__ void sweep3d_tile(const T *__restrict__ src, T *__restrict__ dst,
                                             const JacobiArgs &A, int tx, int ty, int zc,
                                             bool want_norm, double &nacc) {
    constexpr int VX = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    const int x = tx * 64 * VX + lane * VX;
    const int y0 = ty * RY;
    const int zb = 1 + zc * A.chunk_len;
    const int ze = min(zb + A.chunk_len, nz - 1);
    if (zb >= ze) return;
    const int xe = (lane == 0) ? x - 1 : x + VX;
    const bool has_edge = (lane == 0 || lane == 63) && xe >= 0 && xe < nx;
    const T c = stencil_coeff3d<T>();
    const bool periodic = A.bc == CFM_BC_PERIODIC;

    T prev[RY][VX];
    T cur[RY + 2][VX], nxt[RY + 2][VX];
    T ecur[RY], enxt[RY];

    // plane zb-1 (own rows) and zb (own rows + halo rows + edges)
    {
        const T *pp = src + (size_t)(zb - 1) * plane;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            load_row<T, VX, ALIGNED>(pp + (size_t)y * nx, x, nx, y < ny, prev[r]);
        }
        const T *pc = src + (size_t)zb * plane;
#pragma unroll
        for (int i = 0; i < RY + 2; ++i) {
            const int y = y0 - 1 + i;
            const bool ok = y >= 0 && y < ny;
            load_row<T, VX, ALIGNED>(pc + (size_t)(ok ? y : 0) * nx, x, nx, ok, cur[i]);
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            ecur[r] = (has_edge && y < ny) ? pc[(size_t)y * nx + xe] : T(0);
        }
    }

    for (int z = zb; z < ze; ++z) {
        // prefetch plane z+1 (always < nz)
        const T *pn = src + (size_t)(z + 1) * plane;
#pragma unroll
        for (int i = 0; i < RY + 2; ++i) {
            const int y = y0 - 1 + i;
            const bool ok = y >= 0 && y < ny;
            load_row<T, VX, ALIGNED>(pn + (size_t)(ok ? y : 0) * nx, x, nx, ok, nxt[i]);
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            enxt[r] = (has_edge && y < ny) ? pn[(size_t)y * nx + xe] : T(0);
        }

        T *dz = dst + (size_t)z * plane;
        T *dcopy = nullptr;  // periodic full-plane copy target
        if (periodic) {
            if (z == nz - 2) dcopy = dst;                            // -> plane 0
            else if (z == 1) dcopy = dst + (size_t)(nz - 1) * plane;  // -> plane nz-1
        }
        T *dcopy2 = (periodic && z == 1 && nz - 2 == 1) ? dst + (size_t)(nz - 1) * plane : nullptr;

#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            const T *cr = cur[r + 1];
            // lane neighbours (executed by the whole wavefront)
            T west = __shfl_up(cr[VX - 1], 1, 64);
            T east = __shfl_down(cr[0], 1, 64);
            if (lane == 0) west = ecur[r];
            if (lane == 63) east = ecur[r];
            T out[VX];
            bool inter[VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const T E = (e < VX - 1) ? cr[e + 1] : east;
                const T W = (e > 0) ? cr[e - 1] : west;
                const T v = c * (E + W + cur[r + 2][e] + cur[r][e] + nxt[r + 1][e] + prev[r][e]);
                const int xx = x + e;
                inter[e] = xx >= 1 && xx <= nx - 2 && y >= 1 && y <= ny - 2;
                out[e] = inter[e] ? v : cr[e];
                if (want_norm && inter[e]) {
                    const double d = (double)v - (double)cr[e];
                    nacc += d * d;
                }
            }
            if (y >= 1 && y <= ny - 2) {
                T *drow = dz + (size_t)y * nx;
                if (ALIGNED && x >= 1 && x + VX - 1 <= nx - 2) {
                    typename Vec16<T>::type t;
                    T *tp = reinterpret_cast<T *>(&t);
#pragma unroll
                    for (int e = 0; e < VX; ++e) tp[e] = out[e];
                    *reinterpret_cast<typename Vec16<T>::type *>(drow + x) = t;
                } else {
#pragma unroll
                    for (int e = 0; e < VX; ++e)
                        if (inter[e]) drow[x + e] = out[e];
                }
            }
            if (dcopy && y < ny) {
                T *crow = dcopy + (size_t)y * nx;
#pragma unroll
                for (int e = 0; e < VX; ++e)
                    if (x + e < nx) crow[x + e] = out[e];
                if (dcopy2) {
                    T *crow2 = dcopy2 + (size_t)y * nx;
#pragma unroll
                    for (int e = 0; e < VX; ++e)
                        if (x + e < nx) crow2[x + e] = out[e];
                }
            }
        }

        // rotate the register queue
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
            for (int e = 0; e < VX; ++e) prev[r][e] = cur[r + 1][e];
#pragma unroll
        for (int i = 0; i < RY + 2; ++i)
#pragma unroll
            for (int e = 0; e < VX; ++e) cur[i][e] = nxt[i][e];
#pragma unroll
        for (int r = 0; r < RY; ++r) ecur[r] = enxt[r];
    }
}

template <typename T, int RY, bool ALIGNED>
__global__ void __launch_bounds__(256) jacobi3d_kernel(JacobiArgs A) {
    __shared__ double lds4[4];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int ntiles = A.ntx * A.nty;
    const int nunits = ntiles * A.chunks;
    unsigned long long epoch = 0;
    for (int k = 0; k < A.niter; ++k) {
        const int it = A.it0 + k;
        const bool want = A.norm_every > 0 && (it % A.norm_every) == 0;
        if (!A.no_compute) {
            const T *src = static_cast<const T *>((it & 1) ? A.buf1 : A.buf0);
            T *dst = static_cast<T *>((it & 1) ? A.buf0 : A.buf1);
            double nacc = 0.0;
            for (int u = wave_id; u < nunits; u += nwaves) {
                const int zc = u / ntiles;
                const int rem = u - zc * ntiles;
                const int ty = rem / A.ntx;
                const int tx = rem - ty * A.ntx;
                sweep3d_tile<T, RY, ALIGNED>(src, dst, A, tx, ty, zc, want, nacc);
            }
            if (want) {
                const double s = block256_sum(nacc, lds4);
                if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        if (k + 1 < A.niter) {
            if (!grid_barrier(A.sync, epoch)) return;
        }
    }
}

// ------------------------------------------------------------------ 2D ----
template <typename T, int RB, bool ALIGNED>
__device__ __forceinline__ void sweep2d_tile(const T *__restrict__ src, T *__restrict__ dst,
                                             const JacobiArgs &A, int tx, int yc, bool want_norm,
                                             double &nacc) {
    constexpr int VX = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
    const int nx = A.nx, ny = A.ny;
    const int x = tx * 64 * VX + lane * VX;
    const int yb = 1 + yc * A.chunk_len;
    const int ye = min(yb + A.chunk_len, ny - 1);
    if (yb >= ye) return;
    const int xe = (lane == 0) ? x - 1 : x + VX;
    const bool has_edge = (lane == 0 || lane == 63) && xe >= 0 && xe < nx;
    const bool periodic = A.bc == CFM_BC_PERIODIC;
    const T quarter = T(0.25);

    T prow[VX];
    T cur[RB][VX], nxt[RB][VX];
    T ecur[RB], enxt[RB];

    load_row<T, VX, ALIGNED>(src + (size_t)(yb - 1) * nx, x, nx, true, prow);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int y = yb + r;
        const bool ok = y <= ye;  // row ye is the north neighbour of the last row
        load_row<T, VX, ALIGNED>(src + (size_t)(ok ? y : 0) * nx, x, nx, ok, cur[r]);
        ecur[r] = (has_edge && ok) ? src[(size_t)y * nx + xe] : T(0);
    }

    for (int y = yb; y < ye; y += RB) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int yy = y + RB + r;
            const bool ok = yy <= ye;
            load_row<T, VX, ALIGNED>(src + (size_t)(ok ? yy : 0) * nx, x, nx, ok, nxt[r]);
            enxt[r] = (has_edge && ok) ? src[(size_t)yy * nx + xe] : T(0);
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int yy = y + r;
            if (yy < ye) {  // uniform across the wavefront
                const T *cr = cur[r];
                T west = __shfl_up(cr[VX - 1], 1, 64);
                T east = __shfl_down(cr[0], 1, 64);
                if (lane == 0) west = ecur[r];
                if (lane == 63) east = ecur[r];
                T out[VX];
                bool inter[VX];
#pragma unroll
                for (int e = 0; e < VX; ++e) {
                    const T E = (e < VX - 1) ? cr[e + 1] : east;
                    const T W = (e > 0) ? cr[e - 1] : west;
                    const T N = (r + 1 < RB) ? cur[r + 1][e] : nxt[0][e];
                    const T S = (r > 0) ? cur[r - 1][e] : prow[e];
                    const T v = quarter * (E + W + N + S);
                    const int xx = x + e;
                    inter[e] = xx >= 1 && xx <= nx - 2;
                    out[e] = inter[e] ? v : cr[e];
                    if (want_norm && inter[e]) {
                        const double d = (double)v - (double)cr[e];
                        nacc += d * d;
                    }
                }
                T *drow = dst + (size_t)yy * nx;
                if (ALIGNED && x >= 1 && x + VX - 1 <= nx - 2) {
                    typename Vec16<T>::type t;
                    T *tp = reinterpret_cast<T *>(&t);
#pragma unroll
                    for (int e = 0; e < VX; ++e) tp[e] = out[e];
                    *reinterpret_cast<typename Vec16<T>::type *>(drow + x) = t;
                } else {
#pragma unroll
                    for (int e = 0; e < VX; ++e)
                        if (inter[e]) drow[x + e] = out[e];
                }
                if (periodic && (yy == 1 || yy == ny - 2)) {
                    // full-row copies: row ny-2 -> row 0, row 1 -> row ny-1
                    if (yy == ny - 2) {
#pragma unroll
                        for (int e = 0; e < VX; ++e)
                            if (x + e < nx) dst[x + e] = out[e];
                    }
                    if (yy == 1) {
                        T *crow = dst + (size_t)(ny - 1) * nx;
#pragma unroll
                        for (int e = 0; e < VX; ++e)
                            if (x + e < nx) crow[x + e] = out[e];
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < VX; ++e) prow[e] = cur[RB - 1][e];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
#pragma unroll
            for (int e = 0; e < VX; ++e) cur[r][e] = nxt[r][e];
            ecur[r] = enxt[r];
        }
    }
}

template <typename T, int RB, bool ALIGNED>
__global__ void __launch_bounds__(256) jacobi2d_kernel(JacobiArgs A) {
    __shared__ double lds4[4];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int nunits = A.ntx * A.chunks;
    unsigned long long epoch = 0;
    for (int k = 0; k < A.niter; ++k) {
        const int it = A.it0 + k;
        const bool want = A.norm_every > 0 && (it % A.norm_every) == 0;
        if (!A.no_compute) {
            const T *src = static_cast<const T *>((it & 1) ? A.buf1 : A.buf0);
            T *dst = static_cast<T *>((it & 1) ? A.buf0 : A.buf1);
            double nacc = 0.0;
            for (int u = wave_id; u < nunits; u += nwaves) {
                const int yc = u / A.ntx;
                const int tx = u - yc * A.ntx;
                sweep2d_tile<T, RB, ALIGNED>(src, dst, A, tx, yc, want, nacc);
            }
            if (want) {
                const double s = block256_sum(nacc, lds4);
                if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        if (k + 1 < A.niter) {
            if (!grid_barrier(A.sync, epoch)) return;
        }
    }
}

// ------------------------------------------------------------- dispatch ---
using KernelFn = void (*)(JacobiArgs);

template <typename T>
static KernelFn pick_kernel(bool is3d, int rows, bool aligned) {
#define CFM_J3(R) (aligned ? (KernelFn)jacobi3d_kernel<T, R, true> : (KernelFn)jacobi3d_kernel<T, R, false>)
#define CFM_J2(R) (aligned ? (KernelFn)jacobi2d_kernel<T, R, true> : (KernelFn)jacobi2d_kernel<T, R, false>)
    if (is3d) {
        switch (rows) {
            case 1: return CFM_J3(1);
            case 2: return CFM_J3(2);
            case 4: return CFM_J3(4);
            case 8: return CFM_J3(8);
            default: return nullptr;
        }
    }
    switch (rows) {
        case 4: return CFM_J2(4);
        case 8: return CFM_J2(8);
        case 16: return CFM_J2(16);
        default: return nullptr;
    }
#undef CFM_J3
#undef CFM_J2
}

cfm_status current_device_props(DeviceProps *out) {
    int dev = -1;
    CFM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CFM_HIP(hipGetDeviceProperties(&prop, dev));
    out->device = dev;
    out->cus = prop.multiProcessorCount;
    out->coop = prop.cooperativeLaunch;
    std::snprintf(out->name, sizeof(out->name), "%s", prop.gcnArchName);
    return CFM_OK;
}

}  // namespace cfm

using namespace cfm;

struct cfm_jacobi_plan {
    cfm_jacobi_params p{};
    DeviceProps dev{};
    size_t elems = 0;
    size_t elem_size = 0;
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;  // buffer holding the current field
    GridSync *sync = nullptr;
    double *norm_part = nullptr;
    int norm_slots = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    KernelFn kernel = nullptr;
    int grid = 0;
    int chunks = 0, chunk_len = 0, ntx = 0, nty = 0, rows = 0;
    unsigned long long spin_ticks = 0;
};

static cfm_status validate(const cfm_jacobi_params *p) {
    if (!p) return CFM_EINVAL;
    if (p->nx < 3 || p->ny < 3 || p->nz < 1 || p->nz == 2) return CFM_EINVAL;
    if (p->niter < 0) return CFM_EINVAL;
    if (p->prec != CFM_F64 && p->prec != CFM_F32) return CFM_EINVAL;
    if (p->bc != CFM_BC_PERIODIC && p->bc != CFM_BC_DIRICHLET) return CFM_EINVAL;
    if (p->variant < CFM_JACOBI_LAUNCH_PER_ITER || p->variant > CFM_JACOBI_NO_COMPUTE)
        return CFM_EINVAL;
    // 32-bit plane/row arithmetic inside the kernels
    const double elems = (double)p->nx * p->ny * p->nz;
    if (elems > 9.0e15 || (double)p->nx * p->ny > 2.0e9) return CFM_EINVAL;
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_create(const cfm_jacobi_params *p, cfm_jacobi_plan **out) {
    if (!out) return CFM_EINVAL;
    *out = nullptr;
    cfm_status st = validate(p);
    if (st != CFM_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CFM_EDEVICE;
    auto *pl = new (std::nothrow) cfm_jacobi_plan();
    if (!pl) return CFM_ENOMEM;
    pl->p = *p;
    st = current_device_props(&pl->dev);
    if (st != CFM_OK) {
        delete pl;
        return st;
    }
    const bool is3d = p->nz >= 3;
    const bool f64 = p->prec == CFM_F64;
    pl->elem_size = f64 ? 8 : 4;
    pl->elems = (size_t)p->nx * p->ny * (size_t)p->nz;
    const int vx = 16 / (int)pl->elem_size;
    const bool aligned = ((size_t)p->nx * pl->elem_size) % 16 == 0;
    pl->rows = p->tile_rows > 0 ? p->tile_rows : (is3d ? 4 : 8);
    pl->kernel = f64 ? pick_kernel<double>(is3d, pl->rows, aligned)
                     : pick_kernel<float>(is3d, pl->rows, aligned);
    if (!pl->kernel) {
        delete pl;
        return CFM_EINVAL;
    }
    // grid: all co-resident workgroups of 256 threads
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)pl->kernel, 256, 0) !=
            hipSuccess ||
        per_cu < 1) {
        delete pl;
        return CFM_EDEVICE;
    }
    const int max_grid = per_cu * pl->dev.cus;
    pl->grid = p->grid_blocks > 0 ? std::min(p->grid_blocks, max_grid) : max_grid;
    const int nwaves = pl->grid * 4;
    pl->ntx = (p->nx + 64 * vx - 1) / (64 * vx);
    const int interior = is3d ? p->nz - 2 : p->ny - 2;
    if (is3d) {
        pl->nty = (p->ny + pl->rows - 1) / pl->rows;
    } else {
        pl->nty = 1;
    }
    const int tiles = pl->ntx * pl->nty;
    int chunks = p->chunks;
    if (chunks <= 0) {
        chunks = std::max(1, nwaves / tiles);
        const int min_len = is3d ? 16 : 64;
        chunks = std::min(chunks, std::max(1, interior / min_len));
    }
    chunks = std::min(chunks, interior);
    pl->chunk_len = (interior + chunks - 1) / chunks;
    pl->chunks = (interior + pl->chunk_len - 1) / pl->chunk_len;
    pl->norm_slots = p->norm_every > 0 ? (p->niter + p->norm_every - 1) / p->norm_every : 0;
    pl->spin_ticks = spin_ticks(p->spin_limit_us);
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, pl->dev.device) ==
            hipSuccess &&
        rate_khz > 0) {
        const unsigned long long us = p->spin_limit_us > 0 ? p->spin_limit_us : 10000000ull;
        pl->spin_ticks = us * (unsigned long long)rate_khz / 1000ull;
    }

    auto fail = [&](cfm_status s) {
        cfm_jacobi_plan_destroy(pl);
        return s;
    };
    if (hipStreamCreate(&pl->stream) != hipSuccess) return fail(CFM_EDEVICE);
    if (hipEventCreate(&pl->ev0) != hipSuccess || hipEventCreate(&pl->ev1) != hipSuccess)
        return fail(CFM_EDEVICE);
    for (int i = 0; i < 2; ++i)
        if (hipMalloc(&pl->buf[i], pl->elems * pl->elem_size) != hipSuccess)
            return fail(CFM_ENOMEM);
    if (hipMalloc((void **)&pl->sync, sizeof(GridSync)) != hipSuccess) return fail(CFM_ENOMEM);
    if (pl->norm_slots > 0 &&
        hipMalloc((void **)&pl->norm_part, sizeof(double) * pl->norm_slots * pl->grid) != hipSuccess)
        return fail(CFM_ENOMEM);
    *out = pl;
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_load(cfm_jacobi_plan *pl, const void *init_host) {
    if (!pl || !init_host) return CFM_EINVAL;
    const size_t bytes = pl->elems * pl->elem_size;
    CFM_HIP(hipMemcpyAsync(pl->buf[0], init_host, bytes, hipMemcpyHostToDevice, pl->stream));
    CFM_HIP(hipMemcpyAsync(pl->buf[1], pl->buf[0], bytes, hipMemcpyDeviceToDevice, pl->stream));
    CFM_HIP(hipStreamSynchronize(pl->stream));
    pl->cur = 0;
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_run(cfm_jacobi_plan *pl, double *norms, int norm_cap,
                                          int *norm_count, cfm_timing *t) {
    if (!pl) return CFM_EINVAL;
    const auto w0 = std::chrono::steady_clock::now();
    const cfm_jacobi_params &p = pl->p;
    JacobiArgs A{};
    A.buf0 = pl->buf[pl->cur];
    A.buf1 = pl->buf[pl->cur ^ 1];
    A.nx = p.nx;
    A.ny = p.ny;
    A.nz = p.nz;
    A.bc = p.bc;
    A.norm_every = p.norm_every;
    A.chunks = pl->chunks;
    A.chunk_len = pl->chunk_len;
    A.ntx = pl->ntx;
    A.nty = pl->nty;
    A.norm_part = pl->norm_part;
    A.sync = pl->sync;
    A.no_compute = p.variant == CFM_JACOBI_NO_COMPUTE;
    hipLaunchKernelGGL(reset_sync_kernel, dim3(1), dim3(64), 0, pl->stream, pl->sync, pl->spin_ticks);
    CFM_HIP(hipGetLastError());
    CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
    if (p.niter > 0) {
        if (p.variant == CFM_JACOBI_LAUNCH_PER_ITER) {
            for (int k = 0; k < p.niter; ++k) {
                A.niter = 1;
                A.it0 = k;
                hipLaunchKernelGGL(pl->kernel, dim3(pl->grid), dim3(256), 0, pl->stream, A);
                CFM_HIP(hipGetLastError());
            }
        } else {
            A.niter = p.niter;
            A.it0 = 0;
            void *args[] = {&A};
            CFM_HIP(hipLaunchCooperativeKernel((const void *)pl->kernel, dim3(pl->grid), dim3(256),
                                               args, 0, pl->stream));
        }
    }
    CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
    CFM_HIP(hipEventSynchronize(pl->ev1));
    float ms = 0.f;
    CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
    GridSync hs{};
    CFM_HIP(hipMemcpy(&hs, pl->sync, sizeof(hs), hipMemcpyDeviceToHost));
    if (hs.abort) return CFM_ETIMEOUT;
    int produced = 0;
    if (pl->norm_slots > 0 && p.variant != CFM_JACOBI_NO_COMPUTE) {
        std::vector<double> part((size_t)pl->norm_slots * pl->grid);
        CFM_HIP(hipMemcpy(part.data(), pl->norm_part, sizeof(double) * part.size(),
                          hipMemcpyDeviceToHost));
        for (int s = 0; s < pl->norm_slots; ++s) {
            double acc = 0.0;
            for (int b = 0; b < pl->grid; ++b) acc += part[(size_t)s * pl->grid + b];
            if (norms && s < norm_cap) norms[s] = std::sqrt(acc);
        }
        produced = pl->norm_slots;
    }
    if (norm_count) *norm_count = produced;
    if (p.variant != CFM_JACOBI_NO_COMPUTE) pl->cur ^= (p.niter & 1);
    if (t) {
        t->kernel_ms = ms;
        t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        t->grid_blocks = pl->grid;
        t->block_threads = 256;
        t->iterations = p.niter;
        t->outer_iterations = 0;
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_download(cfm_jacobi_plan *pl, void *out_host) {
    if (!pl || !out_host) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(out_host, pl->buf[pl->cur], pl->elems * pl->elem_size, hipMemcpyDeviceToHost));
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_field(cfm_jacobi_plan *pl, void **device_ptr) {
    if (!pl || !device_ptr) return CFM_EINVAL;
    *device_ptr = pl->buf[pl->cur];
    return CFM_OK;
}

extern "C" void cfm_jacobi_plan_destroy(cfm_jacobi_plan *pl) {
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    for (int i = 0; i < 2; ++i)
        if (pl->buf[i]) (void)hipFree(pl->buf[i]);
    if (pl->sync) (void)hipFree(pl->sync);
    if (pl->norm_part) (void)hipFree(pl->norm_part);
    if (pl->ev0) (void)hipEventDestroy(pl->ev0);
    if (pl->ev1) (void)hipEventDestroy(pl->ev1);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

extern "C" cfm_status cfm_jacobi_run(const cfm_jacobi_params *p, const void *init_host,
                                     void *out_host, double *norms, int norm_cap, int *norm_count,
                                     cfm_timing *t) {
    if (!init_host || !out_host) return CFM_EINVAL;
    cfm_jacobi_plan *pl = nullptr;
    cfm_status st = cfm_jacobi_plan_create(p, &pl);
    if (st != CFM_OK) return st;
    st = cfm_jacobi_plan_load(pl, init_host);
    if (st == CFM_OK) st = cfm_jacobi_plan_run(pl, norms, norm_cap, norm_count, t);
    if (st == CFM_OK) st = cfm_jacobi_plan_download(pl, out_host);
    cfm_jacobi_plan_destroy(pl);
    return st;
}
