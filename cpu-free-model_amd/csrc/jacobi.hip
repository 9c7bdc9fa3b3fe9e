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

#include <type_traits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "jacobi_host.hpp"

namespace cfm {

__global__ void reset_sync_kernel(GridSync *s, unsigned long long limit_ticks, int grid) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kGoCopies * 32) s->go[t] = 0;
    if (t == 0) {
        s->abort = 0;
        s->spin_limit_ticks = limit_ticks;
    }
    for (int b = t; b < grid; b += gridDim.x * blockDim.x) s->arrive[b] = 0;
}

// ------------------------------------------------------------ pass clock --
// Phase timers of the Jacobi kernels (TIMED builds, A.timers set): the time
// between barriers goes to the sweeps, the wait at the device-wide barrier to
// the barrier.
template <bool TIMED>
__device__ __forceinline__ bool pass_barrier(const JacobiArgs &A, unsigned int &epoch, PhaseClock<TIMED> &clk) {
    clk.mark(CFM_PHASE_SPMV);
    const bool ok = grid_barrier(A.sync, epoch);
    clk.mark(CFM_PHASE_BARRIER);
    return ok;
}

typedef double v2f64 __attribute__((ext_vector_type(2)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
    using type = v2f64;
};
template <>
struct Vec16<float> {
    using type = v4f32;
};

// VX consecutive elements as 16-byte vectors (p 16-byte aligned).
template <typename T, int VX>
__device__ __forceinline__ void ld_vec(const T *p, T (&v)[VX]) {
    constexpr int E = 16 / sizeof(T);
    static_assert(VX % E == 0, "whole 16-byte vectors");
#pragma unroll
    for (int q = 0; q < VX / E; ++q) {
        const typename Vec16<T>::type t = *reinterpret_cast<const typename Vec16<T>::type *>(p + q * E);
#pragma unroll
        for (int e = 0; e < E; ++e) v[q * E + e] = t[e];
    }
}
template <typename T, int VX>
__device__ __forceinline__ void st_vec(T *p, const T (&v)[VX]) {
    constexpr int E = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < VX / E; ++q) {
        typename Vec16<T>::type t;
#pragma unroll
        for (int e = 0; e < E; ++e) t[e] = v[q * E + e];
        *reinterpret_cast<typename Vec16<T>::type *>(p + q * E) = t;
    }
}

// x-neighbour exchange inside a wavefront: lane i receives lane i-1's value
// (lane_up1) or lane i+1's (lane_dn1); lane 0 / lane 63 receive their own
// value and are overwritten by the caller.  8-byte values move with DPP
// whole-wavefront shifts (two v_mov_b32_dpp and no LDS traffic, instead of
// two ds_bpermute): FP64 two-sweep kernels 469 -> 478 GLUP/s (3D 512^3) and
// 537 -> 617 (2D 16384^2); 4-byte values measured no better, they keep
// __shfl, as do the one-sweep kernels (DPP measured no faster there).
template <typename T>
__device__ __forceinline__ T lane_up1(T v) {
    if constexpr (sizeof(T) == 8) {
        union {
            T t;
            int i[2];
        } a, r;
        a.t = v;
        r.i[0] = __builtin_amdgcn_update_dpp(a.i[0], a.i[0], 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
        r.i[1] = __builtin_amdgcn_update_dpp(a.i[1], a.i[1], 0x138, 0xf, 0xf, false);
        return r.t;
    } else {
        return __shfl_up(v, 1, 64);
    }
}
template <typename T>
__device__ __forceinline__ T lane_dn1(T v) {
    if constexpr (sizeof(T) == 8) {
        union {
            T t;
            int i[2];
        } a, r;
        a.t = v;
        r.i[0] = __builtin_amdgcn_update_dpp(a.i[0], a.i[0], 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
        r.i[1] = __builtin_amdgcn_update_dpp(a.i[1], a.i[1], 0x130, 0xf, 0xf, false);
        return r.t;
    } else {
        return __shfl_down(v, 1, 64);
    }
}

// The same shifts with zeros shifted in at lane 0 / lane 63 (bound_ctrl): no copy
// of the old value into the destination first, for callers that replace those
// two lanes anyway.
template <typename T>
__device__ __forceinline__ T lane_up1z(T v) {
    static_assert(sizeof(T) == 8 || sizeof(T) == 4, "4- or 8-byte values");
    if constexpr (sizeof(T) == 8) {
        union {
            T t;
            int i[2];
        } a, r;
        a.t = v;
        r.i[0] = __builtin_amdgcn_mov_dpp(a.i[0], 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
        r.i[1] = __builtin_amdgcn_mov_dpp(a.i[1], 0x138, 0xf, 0xf, true);
        return r.t;
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
    }
}
template <typename T>
__device__ __forceinline__ T lane_dn1z(T v) {
    static_assert(sizeof(T) == 8 || sizeof(T) == 4, "4- or 8-byte values");
    if constexpr (sizeof(T) == 8) {
        union {
            T t;
            int i[2];
        } a, r;
        a.t = v;
        r.i[0] = __builtin_amdgcn_mov_dpp(a.i[0], 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
        r.i[1] = __builtin_amdgcn_mov_dpp(a.i[1], 0x130, 0xf, 0xf, true);
        return r.t;
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
    }
}

// Interior stores of the new field: it is not read again before the next
// sweep, by which time it has left the caches, so it may bypass them.
#ifndef CFM_JACOBI_NT_STORE
#define CFM_JACOBI_NT_STORE 1
#endif
template <typename V>
__device__ __forceinline__ void store16(V *p, V v) {
#if CFM_JACOBI_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

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
            ld_vec<T, VX>(p + x, v);
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
#pragma unroll
                    for (int e = 0; e < VX; ++e) t[e] = out[e];
                    store16(reinterpret_cast<typename Vec16<T>::type *>(drow + x), t);
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

template <typename T, int RY, bool ALIGNED, int BLOCK, bool TIMED>
__global__ void __launch_bounds__(BLOCK) jacobi3d_kernel(JacobiArgs A) {
    __shared__ double lds[BLOCK / 64];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int ntiles = A.ntx * A.nty;
    const int nunits = ntiles * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
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
                const double s = block_sum_waves<BLOCK / 64>(nacc, lds);
                if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        if (k + 1 < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// ------------------------------------------------------- 3D, workgroup tile --
// A workgroup of WX*WY wavefronts owns a tile of TW = WX*64*VX columns by
// TH = WY*RY rows and streams it through a z-chunk.  Each wavefront keeps
// only its own RY rows of planes z-1, z, z+1 in registers; the rows just
// outside its own (the neighbour wavefront's first/last row, or the tile's
// halo row loaded from global memory by the edge wavefronts) and the x-edge
// values at its lanes 0/63 are exchanged through LDS, double-buffered by the
// plane parity, one __syncthreads per plane.  Global loads per plane: the
// tile's own rows once, plus two halo rows and two edge columns per tile.
template <typename T, int RY, int WX, int WY>
struct WgTile {
    static constexpr int VX = 16 / sizeof(T);
    static constexpr int SEG = 64 * VX;
    static constexpr int TW = SEG * WX;
    static constexpr int TH = RY * WY;
    static constexpr int BLOCK = 64 * WX * WY;
    // per plane parity
    struct Plane {
        T first[WY + 1][TW];    // first row of wave k (k < WY); [WY] = halo row y0 + TH
        T last[WY + 1][TW];     // [0] = halo row y0 - 1; [k] = last row of wave k-1
        T edge_l[WY][WX + 1][RY];  // [k] = first column of wave k; [WX] = column x0 + TW
        T edge_r[WY][WX + 1][RY];  // [0] = column x0 - 1; [k] = last column of wave k-1
    };
};

template <typename T, int RY, int WX, int WY, bool ALIGNED>
__device__ __forceinline__ void sweep3d_wgtile(const T *__restrict__ src, T *__restrict__ dst,
                                               const JacobiArgs &A, int tx, int ty, int zc,
                                               bool want_norm, double &nacc,
                                               typename WgTile<T, RY, WX, WY>::Plane *lds) {
    using Tl = WgTile<T, RY, WX, WY>;
    constexpr int VX = Tl::VX;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wx = wave % WX, wy = wave / WX;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    const int xl = wx * Tl::SEG + lane * VX;  // column inside the tile
    const int x0 = tx * Tl::TW;
    const int x = x0 + xl;
    const int ty0 = ty * Tl::TH;
    const int y0 = ty0 + wy * RY;
    const int zb = 1 + zc * A.chunk_len;
    const int ze = min(zb + A.chunk_len, nz - 1);
    if (zb >= ze) return;  // uniform over the workgroup
    const T c = stencil_coeff3d<T>();
    const bool periodic = A.bc == CFM_BC_PERIODIC;
    // global halo duties of this wavefront
    const bool top = wy == 0, bot = wy == WY - 1;
    const int yt = ty0 - 1, yb = ty0 + Tl::TH;
    const bool ld_top = top && yt >= 0;
    const bool ld_bot = bot && yb < ny;
    const bool ld_el = wx == 0 && lane == 0 && x0 - 1 >= 0;             // column x0-1
    const bool ld_er = wx == WX - 1 && lane == 63 && x0 + Tl::TW < nx;  // column x0+TW

    // planes z-1, z, z+1 in use and z+2 in flight: the loads issued at step z
    // are consumed at step z+1, so a whole step covers their latency
    T prev[RY][VX], cur[RY][VX], nxt[RY][VX], nx2[RY][VX];
    T hcur[VX] = {}, hnxt[VX] = {}, hnx2[VX] = {};  // this wave's halo row (top or bottom wave only)
    T ecur[RY] = {}, enxt[RY] = {}, enx2[RY] = {};  // this lane's edge column (x-edge lanes only)

    auto load_plane = [&](int z, T (&rows)[RY][VX], T (&h)[VX], T (&e)[RY]) {
        const T *pz = src + (size_t)z * plane;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            load_row<T, VX, ALIGNED>(pz + (size_t)(y < ny ? y : 0) * nx, x, nx, y < ny, rows[r]);
        }
        if (top || bot) {
            const bool ok = top ? ld_top : ld_bot;
            const int y = top ? yt : yb;
            load_row<T, VX, ALIGNED>(pz + (size_t)(ok ? y : 0) * nx, x, nx, ok, h);
        }
        if (ld_el || ld_er) {
            const int xe = ld_el ? x0 - 1 : x0 + Tl::TW;
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int y = y0 + r;
                e[r] = y < ny ? pz[(size_t)y * nx + xe] : T(0);
            }
        }
    };

    {
        T hd[VX], ed[RY];
        load_plane(zb - 1, prev, hd, ed);
        load_plane(zb, cur, hcur, ecur);
        load_plane(zb + 1, nxt, hnxt, enxt);  // zb+1 <= ze <= nz-1
    }
    for (int z = zb; z < ze; ++z) {
        if (z + 2 <= nz - 1) load_plane(z + 2, nx2, hnx2, enx2);  // uniform over the workgroup
        typename Tl::Plane &P = lds[z & 1];
        // publish this wave's boundary rows / columns of plane z
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            P.first[wy][xl + e] = cur[0][e];
            P.last[wy + 1][xl + e] = cur[RY - 1][e];
        }
        if (top)
#pragma unroll
            for (int e = 0; e < VX; ++e) P.last[0][xl + e] = hcur[e];
        if (bot)
#pragma unroll
            for (int e = 0; e < VX; ++e) P.first[WY][xl + e] = hcur[e];
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < RY; ++r) P.edge_l[wy][wx][r] = cur[r][0];
        if (lane == 63)
#pragma unroll
            for (int r = 0; r < RY; ++r) P.edge_r[wy][wx + 1][r] = cur[r][VX - 1];
        if (wx == 0 && lane == 0)
#pragma unroll
            for (int r = 0; r < RY; ++r) P.edge_r[wy][0][r] = ecur[r];
        if (wx == WX - 1 && lane == 63)
#pragma unroll
            for (int r = 0; r < RY; ++r) P.edge_l[wy][WX][r] = ecur[r];
        __syncthreads();

        T *dz = dst + (size_t)z * plane;
        T *dcopy = nullptr;
        if (periodic) {
            if (z == nz - 2) dcopy = dst;
            else if (z == 1) dcopy = dst + (size_t)(nz - 1) * plane;
        }
        T *dcopy2 = (periodic && z == 1 && nz - 2 == 1) ? dst + (size_t)(nz - 1) * plane : nullptr;

        T north[VX], south[VX];
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            south[e] = P.last[wy][xl + e];
            north[e] = P.first[wy + 1][xl + e];
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int y = y0 + r;
            const T *cr = cur[r];
            T west = __shfl_up(cr[VX - 1], 1, 64);
            T east = __shfl_down(cr[0], 1, 64);
            if (lane == 0) west = P.edge_r[wy][wx][r];
            if (lane == 63) east = P.edge_l[wy][wx + 1][r];
            T out[VX];
            bool inter[VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const T E = (e < VX - 1) ? cr[e + 1] : east;
                const T W = (e > 0) ? cr[e - 1] : west;
                const T N = (r + 1 < RY) ? cur[r + 1][e] : north[e];
                const T S = (r > 0) ? cur[r - 1][e] : south[e];
                const T v = c * (E + W + N + S + nxt[r][e] + prev[r][e]);
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
#pragma unroll
                    for (int e = 0; e < VX; ++e) t[e] = out[e];
                    store16(reinterpret_cast<typename Vec16<T>::type *>(drow + x), t);
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
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                prev[r][e] = cur[r][e];
                cur[r][e] = nxt[r][e];
                nxt[r][e] = nx2[r][e];
            }
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            hcur[e] = hnxt[e];
            hnxt[e] = hnx2[e];
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            ecur[r] = enxt[r];
            enxt[r] = enx2[r];
        }
    }
    __syncthreads();  // the LDS planes are reused by the next unit
}

template <typename T, int RY, int WX, int WY, bool ALIGNED, bool TIMED>
__global__ void __launch_bounds__(64 * WX * WY) jacobi3d_wg_kernel(JacobiArgs A) {
    using Tl = WgTile<T, RY, WX, WY>;
    __shared__ typename Tl::Plane planes[2];
    __shared__ double lds[Tl::BLOCK / 64];
    const int g = gridDim.x, b = blockIdx.x;
    const int lb = (g & 7) == 0 ? (b & 7) * (g >> 3) + (b >> 3) : b;  // XCD-contiguous order
    const int ntiles = A.ntx * A.nty;
    const int nunits = ntiles * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    for (int k = 0; k < A.niter; ++k) {
        const int it = A.it0 + k;
        const bool want = A.norm_every > 0 && (it % A.norm_every) == 0;
        if (!A.no_compute) {
            const T *src = static_cast<const T *>((it & 1) ? A.buf1 : A.buf0);
            T *dst = static_cast<T *>((it & 1) ? A.buf0 : A.buf1);
            double nacc = 0.0;
            for (int u = lb; u < nunits; u += g) {
                const int zc = u / ntiles;
                const int rem = u - zc * ntiles;
                const int ty = rem / A.ntx;
                const int tx = rem - ty * A.ntx;
                sweep3d_wgtile<T, RY, WX, WY, ALIGNED>(src, dst, A, tx, ty, zc, want, nacc, planes);
            }
            if (want) {
                const double s = block_sum_waves<Tl::BLOCK / 64>(nacc, lds);
                if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        if (k + 1 < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// ------------------------------------------------- 3D, two sweeps per pass --
// Temporal blocking for the Dirichlet case when one tile spans the x extent
// (nx <= TW): a pass reads the field at sweep t once and writes sweep t+2,
// so HBM sees one read and one write per two updates.  The intermediate
// sweep t+1 ("level 1") lives only in registers and LDS.
//
// Workgroup = WX x WY wavefronts; wave (wx, wy) owns columns
// [wx*SEG, wx*SEG + SEG) of row y = ty*WY + wy and streams planes p of a
// z-chunk.  At step p it computes level 1 on plane p from level-0 planes
// p-1, p, p+1, publishes its level-1 row and x-edge values to LDS, and
// computes level 2 on plane p-1 from level-1 planes p-2, p-1, p and the
// y-neighbour rows of the waves above and below, which the other waves
// published at step p-1: they are read at the start of step p, so their LDS
// latency overlaps level 1, and the step's one __syncthreads comes last
// (1024-thread builds, with half the registers per wave, read them after that
// barrier, placed between the two levels).
// The tile's top and bottom waves also compute level 1 on the row just
// outside the tile (redundantly with the neighbouring tile), and every chunk
// recomputes level 1 on the plane just outside each end.  Every level-1 value
// comes from the same expression as a single sweep and boundary points keep
// their value, so the result is bit-identical to two single sweeps.
//
// Level-0 rows of a wave: "column" rows are needed on planes p-1, p, p+1
// (its own row, and the tile-halo row of the top/bottom wave) and are
// prefetched CLA planes ahead -- they are the HBM stream; "side" rows are
// needed on plane p only (the y-neighbours of the level-1 rows) and are
// fetched one plane ahead (the neighbouring wave or tile has streamed them
// into L2).  x-edge values of the column rows come from the neighbouring
// waves through LDS.  The pass is bound by instruction issue rather than by
// HBM, so the per-row work is kept free of per-lane selects and masks.  LDS is triple-buffered by plane (written
// at step p, read at step p+1, rewritten at step p+3).
template <typename T, int WX, int WY, int VXT = 16 / sizeof(T)>
struct T2Tile {
    static constexpr int VX = VXT;  // elements per lane and row (whole 16-byte vectors)
    static constexpr int SEG = 64 * VX;
    static constexpr int TW = SEG * WX;
    static constexpr int BLOCK = 64 * WX * WY;
    // level-2 rows read before level 1 (their registers live across it): up to 512 threads, where a
    // wave has 256 VGPRs; the 1024-thread builds (128 VGPRs) read them after the step's barrier
    static constexpr bool HOIST = BLOCK <= 512;
    struct Buf {
        T row[WY][TW];         // level-1 own row of each wave (plane p)
        T l1e[WY][WX][2];      // its first (lane 0) and last (lane 63) element
        T l0e[WY][WX][2][2];   // level-0 column rows of plane p+1: [slot][first, last]
        T l0x[WY][WX][2][64];  // the same edges, one slot per lane (sweep3d_t2r: unmasked stores)
    };
};

// FULL: nx == TW, so every lane's VX columns lie inside the rows (unmasked
// vector loads and stores; the two boundary columns are rewritten with their
// own, unchanged value).  Wave roles (top / middle / bottom of the tile) are
// wave-uniform branches, so no per-lane selects are spent on them.
//
// PER: the reference's periodic z handling (after every sweep the last
// interior plane is copied to plane 0 and the first to plane nz-1).  Level 0
// is read as stored (its planes 0 and nz-1 hold the previous pass's copies,
// or the initial field); level 1 on plane 0 is therefore level 1 on plane
// nz-2, and on plane nz-1 it is level 1 on plane 1: the chunk starting at
// plane 1 computes the former in a prologue from stored planes nz-3..nz-1,
// the chunk ending at plane nz-2 the latter from stored planes 0..2 in its
// last step.  Whoever writes level 2 (or, on a one-sweep pass, level 1) of
// planes nz-2 and 1 also writes the whole row into plane 0 / nz-1.
template <typename T, int WX, int WY, int VXT, bool ALIGNED, bool FULL, bool PER, int CLA = 1>
__device__ __forceinline__ void sweep3d_t2(const T *__restrict__ src, T *__restrict__ dst, const JacobiArgs &A,
                                           int ty, int zc, bool single, bool want0, bool want1, double &n0,
                                           double &n1, typename T2Tile<T, WX, WY, VXT>::Buf *lds) {
    using Tl = T2Tile<T, WX, WY, VXT>;
    constexpr int VX = Tl::VX;
    constexpr bool kHoist = Tl::HOIST;
    static_assert(WY >= 2, "the top and bottom wave must differ");
    static_assert(!FULL || ALIGNED, "full-width rows are 16-byte aligned");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: roles become scalar branches
    const int wx = wave % WX, wy = wave / WX;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    const int x = wx * Tl::SEG + lane * VX;
    const int y = ty * WY + wy;
    const int zb = 1 + zc * A.chunk_len;
    const int ze = min(zb + A.chunk_len, nz - 1);
    if (zb >= ze) return;  // uniform over the workgroup
    const T c = stencil_coeff3d<T>();
    const bool top = wy == 0, bot = wy == WY - 1;
    const bool lane_w = lane == 0, lane_e = lane == 63;
    // column rows: [0] own row, [1] tile-halo row (top: y-1, bottom: y+1)
    const int yc1 = top ? y - 1 : y + 1;
    const bool has_h = top || bot;
    // side rows (plane p only): [0] below the lowest level-1 row, [1] above the highest
    const int ys0 = top ? y - 2 : y - 1;
    const int ys1 = bot ? y + 2 : y + 1;
    bool xin[VX];  // interior columns of this lane
#pragma unroll
    for (int k = 0; k < VX; ++k) xin[k] = x + k >= 1 && x + k <= nx - 2;

    constexpr int NC = 3 + CLA;  // column rows, planes p-1 .. p+1+CLA (the last CLA in flight)
    T col[NC][2][VX];
    T sid[2][2][VX];  // side rows, planes p and p+1 (p+1 in flight)
    T l1o[3][VX];     // level-1 own row, planes p-2, p-1, p
    T l1h[2][VX];     // level-1 tile-halo row (top/bottom waves), planes p-1, p

    auto ld = [&](int z, int yy, bool use, T (&v)[VX]) {
        if (use && z >= 0 && z < nz && yy >= 0 && yy < ny) {  // wave-uniform
            const T *pr = src + (size_t)z * plane + (size_t)yy * nx;
            if (FULL) {
                ld_vec<T, VX>(pr + x, v);
            } else {
                load_row<T, VX, ALIGNED>(pr, x, nx, true, v);
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) v[k] = T(0);
        }
    };
    auto ld_col = [&](int z, T (&v)[2][VX]) {
        ld(z, y, true, v[0]);
        ld(z, yc1, has_h, v[1]);
    };
    auto ld_side = [&](int z, T (&v)[2][VX]) {
        ld(z, ys0, true, v[0]);
        ld(z, ys1, true, v[1]);
    };
    auto put_l0e = [&](typename Tl::Buf &B, const T (&v)[2][VX]) {
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            if (lane_w) B.l0e[wy][wx][sl][0] = v[sl][0];
            if (lane_e) B.l0e[wy][wx][sl][1] = v[sl][VX - 1];
        }
    };
    // one level-1 row: out = stencil where interior, else the centre value
    auto l1row = [&](const typename Tl::Buf &B, int sl, int yy, bool pin, const T (&cr)[VX], const T (&N)[VX],
                     const T (&S)[VX], const T (&U)[VX], const T (&D)[VX], T (&out)[VX], bool norm) {
        T west = lane_up1(cr[VX - 1]);
        T east = lane_dn1(cr[0]);
        if (lane_w) west = wx > 0 ? B.l0e[wy][wx > 0 ? wx - 1 : 0][sl][1] : T(0);
        if (lane_e) east = wx < WX - 1 ? B.l0e[wy][wx < WX - 1 ? wx + 1 : 0][sl][0] : T(0);
        if (pin && yy >= 1 && yy <= ny - 2) {  // wave-uniform
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                const T E = (k < VX - 1) ? cr[k + 1] : east;
                const T W = (k > 0) ? cr[k - 1] : west;
                const T v = c * (E + W + N[k] + S[k] + U[k] + D[k]);
                out[k] = (FULL && k > 0 && k < VX - 1) ? v : (xin[k] ? v : cr[k]);  // FULL: inner columns are interior
                if (norm && xin[k]) {
                    const double d = (double)v - (double)cr[k];
                    n0 += d * d;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) out[k] = cr[k];
        }
    };
    auto store_row = [&](int z, const T (&out)[VX], const bool (&inter)[VX]) {
        T *drow = dst + (size_t)z * plane + (size_t)y * nx;
        constexpr int E = 16 / sizeof(T);
        if (FULL || (ALIGNED && x >= 1 && x + VX - 1 <= nx - 2)) {
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                typename Vec16<T>::type t;
#pragma unroll
                for (int e = 0; e < E; ++e) t[e] = out[q * E + e];
                store16(reinterpret_cast<typename Vec16<T>::type *>(drow + x + q * E), t);
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k)
                if (inter[k]) drow[x + k] = out[k];
        }
    };

    // the whole row (boundary columns included) of a periodic halo plane
    auto store_copy = [&](int z, const T (&out)[VX]) {
        if (y >= ny) return;  // wave-uniform
        T *drow = dst + (size_t)z * plane + (size_t)y * nx;
        if (FULL) {
            constexpr int E = 16 / sizeof(T);
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                typename Vec16<T>::type t;
#pragma unroll
                for (int e = 0; e < E; ++e) t[e] = out[q * E + e];
                store16(reinterpret_cast<typename Vec16<T>::type *>(drow + x + q * E), t);
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k)
                if (x + k < nx) drow[x + k] = out[k];
        }
    };
    // level 1 on the plane whose level-0 rows are in col[0..2] / sid[0]: own row,
    // and the tile-halo row of the top/bottom wave
    auto level1 = [&](const typename Tl::Buf &Bp, bool pin, bool nrm) {
        if (top) {
            l1row(Bp, 0, y, pin, col[1][0], sid[0][1], col[1][1], col[2][0], col[0][0], l1o[2], nrm);
            l1row(Bp, 1, y - 1, pin, col[1][1], col[1][0], sid[0][0], col[2][1], col[0][1], l1h[1], false);
        } else if (bot) {
            l1row(Bp, 0, y, pin, col[1][0], col[1][1], sid[0][0], col[2][0], col[0][0], l1o[2], nrm);
            l1row(Bp, 1, y + 1, pin, col[1][1], sid[0][1], col[1][0], col[2][1], col[0][1], l1h[1], false);
        } else {
            l1row(Bp, 0, y, pin, col[1][0], sid[0][1], sid[0][0], col[2][0], col[0][0], l1o[2], nrm);
        }
    };

    int p0 = single ? zb : zb - 1;
    const int p1 = single ? ze - 1 : ze;  // inclusive
    if (PER && !single && zb == 1) {
        // level 1 on plane 0 = level 1 on plane nz-2, from stored planes nz-3 .. nz-1
        ld_col(nz - 3, col[0]);
        ld_col(nz - 2, col[1]);
        ld_col(nz - 1, col[2]);
        ld_side(nz - 2, sid[0]);
        put_l0e(lds[0], col[1]);
        __syncthreads();
        level1(lds[0], true, false);
#pragma unroll
        for (int k = 0; k < VX; ++k) {
            l1o[1][k] = l1o[2][k];
            l1h[0][k] = l1h[1][k];
        }
        p0 = 1;  // the stream starts at plane 1 (its level-0 edges go to lds[1], not lds[0])
    }
#pragma unroll
    for (int i = 0; i < NC - 1; ++i) ld_col(p0 - 1 + i, col[i]);
    ld_side(p0, sid[0]);
    put_l0e(lds[p0 % 3], col[1]);
    __syncthreads();
    for (int p = p0; p <= p1; ++p) {
        ld_col(p + 1 + CLA, col[NC - 1]);  // consumed at step p+CLA
        ld_side(p + 1, sid[1]);  // consumed at step p+1
        typename Tl::Buf &Bp = lds[p % 3];
        bool pin = p >= 1 && p <= nz - 2;  // interior plane
        if (PER && !single && p == nz - 1) {
            // level 1 on plane nz-1 = level 1 on plane 1, from stored planes 0 .. 2 (nobody
            // reads lds[p % 3] between the previous step's barrier and the one below)
            ld_col(0, col[0]);
            ld_col(1, col[1]);
            ld_col(2, col[2]);
            ld_side(1, sid[0]);
            put_l0e(Bp, col[1]);
            __syncthreads();
            pin = true;
        }
        const bool nrm = want0 && p >= zb && p < ze;
        // level-2 inputs (two-sweep passes): the level-1 rows of plane p-1 and their x edges,
        // published at step p-1 before the barrier that ended it, are read before level 1 of this
        // step so that their LDS latency overlaps its arithmetic
        const int zo = p - 1;
        const bool l2 = !single && zo >= zb && zo >= 1 && zo <= nz - 2 && y >= 1 && y <= ny - 2;  // wave-uniform
        const typename Tl::Buf &R = lds[(p + 2) % 3];  // written at step p-1
        T N[VX], S[VX], wn2 = T(0), en2 = T(0);
        auto l2_reads = [&] {
            if (l2) {
                if (!top) ld_vec<T, VX>(&R.row[wy - 1 < 0 ? 0 : wy - 1][x], S);
                if (!bot) ld_vec<T, VX>(&R.row[wy + 1 >= WY ? WY - 1 : wy + 1][x], N);
                if (lane_w) wn2 = wx > 0 ? R.l1e[wy][wx > 0 ? wx - 1 : 0][1] : T(0);
                if (lane_e) en2 = wx < WX - 1 ? R.l1e[wy][wx < WX - 1 ? wx + 1 : 0][0] : T(0);
            }
        };
        if constexpr (kHoist) l2_reads();
        level1(Bp, pin, nrm);
        typename Tl::Buf &Bn = lds[(p + 1) % 3];
        if (single) {
            // one sweep only: store the own level-1 row of plane p
            if (y >= 1 && y <= ny - 2) store_row(p, l1o[2], xin);
            if (PER) {  // l1o[2] holds the stored value wherever the point is not interior
                if (p == nz - 2) store_copy(0, l1o[2]);
                if (p == 1) store_copy(nz - 1, l1o[2]);
            }
            put_l0e(Bn, col[2]);
            __syncthreads();
        } else {
            // ---- publish the own level-1 row of plane p and the level-0 edges of plane p+1
            st_vec<T, VX>(&Bp.row[wy][x], l1o[2]);
            if (lane_w) Bp.l1e[wy][wx][0] = l1o[2][0];
            if (lane_e) Bp.l1e[wy][wx][1] = l1o[2][VX - 1];
            put_l0e(Bn, col[2]);
            if constexpr (!kHoist) {
                __syncthreads();
                l2_reads();
            }
            // ---- level 2 on plane p-1
            if (l2) {
                if (top) {
#pragma unroll
                    for (int k = 0; k < VX; ++k) S[k] = l1h[0][k];
                }
                if (bot) {
#pragma unroll
                    for (int k = 0; k < VX; ++k) N[k] = l1h[0][k];
                }
                const T *cr = l1o[1];
                T west = lane_up1(cr[VX - 1]);
                T east = lane_dn1(cr[0]);
                if (lane_w) west = wn2;
                if (lane_e) east = en2;
                T out[VX];
#pragma unroll
                for (int k = 0; k < VX; ++k) {
                    const T E = (k < VX - 1) ? cr[k + 1] : east;
                    const T Wv = (k > 0) ? cr[k - 1] : west;
                    const T v = c * (E + Wv + N[k] + S[k] + l1o[2][k] + l1o[0][k]);
                    out[k] = (FULL && k > 0 && k < VX - 1) ? v : (xin[k] ? v : cr[k]);  // FULL: inner columns are interior
                    if (want1 && xin[k]) {
                        const double d = (double)v - (double)cr[k];
                        n1 += d * d;
                    }
                }
                store_row(zo, out, xin);
                if (PER) {
                    if (zo == nz - 2) store_copy(0, out);
                    if (zo == 1) store_copy(nz - 1, out);
                }
            } else if (PER && zo >= zb && (zo == 1 || zo == nz - 2) && (y == 0 || y == ny - 1)) {
                // boundary rows of a copied plane keep their stored values, which level 1 of
                // plane zo (l1o[1]) carries there (col[0] may hold the reloaded plane 0 here)
                if (zo == nz - 2) store_copy(0, l1o[1]);
                if (zo == 1) store_copy(nz - 1, l1o[1]);
            }
            if constexpr (kHoist) __syncthreads();  // this step's rows and edges are published for the next step
        }
        // ---- slide the windows
#pragma unroll
        for (int k = 0; k < VX; ++k) {
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
#pragma unroll
                for (int i = 0; i < NC - 1; ++i) col[i][sl][k] = col[i + 1][sl][k];
                sid[0][sl][k] = sid[1][sl][k];
            }
            l1o[0][k] = l1o[1][k];
            l1o[1][k] = l1o[2][k];
            l1h[0][k] = l1h[1][k];
        }
    }
    __syncthreads();  // the LDS buffers are reused by the next unit
}

// Two sweeps per pass with a two-step load lookahead.  In sweep3d_t2 the
// level-0 window slides by register moves, and moving the register a load is
// still writing forces the wave to wait for that load at the end of the very
// step that issued it, so at most one step of compute covers the memory
// latency whatever the lookahead distance.  Here the column rows of plane
// p+3 land in a two-slot ring whose slot is fixed at compile time (the plane
// loop is unrolled by two), and only the slot issued one step earlier is
// moved into the window: those loads stay in flight across the next step.
// The side rows (plane p+1, mostly L2 hits: the neighbouring wave streamed
// them) are issued first in each step, so waiting for them at its end leaves
// the ring loads issued after them in flight (the load counter is in order).
// Rows come through a buffer resource: a lane's column is a fixed vector
// offset, the row a scalar offset, and rows or planes outside the field get
// an offset beyond the range check, which returns zeros, so every wave issues
// the same unconditional loads each step (the compiler's wait counts stay
// exact).  Same arithmetic, order and results as sweep3d_t2 (16-byte aligned
// rows; fields below 2 GiB).
template <int S>
struct Slot {
    static constexpr int value = S;
};

//
// XT: rows wider than one tile.  Tile tx covers columns X0 .. X0 + TW - 1 with
// X0 = tx (TW - E) (16-byte aligned).  Level 1 is computed on all of them: the
// neighbours of the two outer columns, X0 - 1 and X0 + TW, are level-0 values
// that the edge lanes load themselves, one plane step ahead like the side rows.
// Level 2 is then valid on X0 + 1 .. X0 + TW - 2, and the tile owns (stores and
// adds to the norms) X0 + 1 .. X0 + TW - E, so consecutive tiles own consecutive
// columns; each pass recomputes E + 1 level-1 columns per tile boundary.
//
// WIN: read a unit's planes through a window (fields of 2 GiB and more, and the x-tile builds);
// without it the buffer covers the whole field, which must then be below 2 GiB (the window's
// bookkeeping costs the 512^3 pass 2 %).
template <typename T, int WX, int WY, int VXT, bool FULL, bool PER, bool XT, bool WIN>
__device__ __forceinline__ void sweep3d_t2r(const T *__restrict__ src, T *__restrict__ dst, const JacobiArgs &A,
                                            int tx, int ty, int zc, bool single, bool want0, bool want1, double &n0,
                                            double &n1, typename T2Tile<T, WX, WY, VXT>::Buf *lds) {
    using Tl = T2Tile<T, WX, WY, VXT>;
    constexpr int VX = Tl::VX;
    constexpr bool kHoist = Tl::HOIST;
    constexpr int E = 16 / sizeof(T);
    static_assert(WY >= 2, "the top and bottom wave must differ");
    static_assert(VX % E == 0, "whole 16-byte vectors");
    static_assert(!(XT && FULL), "full-width rows are one tile");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: roles become scalar branches
    const int wx = wave % WX, wy = wave / WX;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    const int X0 = XT ? tx * (Tl::TW - E) : 0;
    const int xl = wx * Tl::SEG + lane * VX;  // the lane's first column in the tile (LDS rows)
    const int x = X0 + xl;                    // ... and in the field
    // owned columns (stores, norms) and, for the periodic halo-plane copies, owned whole-row columns
    const int olo = XT ? max(1, X0 + 1) : 1;
    const int ohi = XT ? min(nx - 2, X0 + Tl::TW - E) : nx - 2;
    const int clo = XT && X0 > 0 ? X0 + 1 : 0;
    const int chi = XT && X0 + Tl::TW - E < nx - 2 ? X0 + Tl::TW - E : nx - 1;
    const int y = ty * WY + wy;
    const int zb = 1 + zc * A.chunk_len;
    const int ze = min(zb + A.chunk_len, nz - 1);
    if (zb >= ze) return;  // uniform over the workgroup
    const T c = stencil_coeff3d<T>();
    const bool top = wy == 0, bot = wy == WY - 1;
    const bool lane_w = lane == 0, lane_e = lane == 63;
    // FP32: the tile's top and bottom wavefronts, which compute two level-1 rows per step and set
    // the pace at every barrier, take the issue slot first (512^3: 1075 -> 1102 GLUP/s, periodic
    // 904 -> 918; FP64 measured 1.4-2.9 % slower with it)
    if constexpr (sizeof(T) == 4)
        if (top || bot) __builtin_amdgcn_s_setprio(1);
    const int yc1 = top ? y - 1 : y + 1;
    const bool has_h = top || bot;
    const int ys0 = top ? y - 2 : y - 1;
    const int ys1 = bot ? y + 2 : y + 1;
    bool xin[VX], own[VX];
#pragma unroll
    for (int k = 0; k < VX; ++k) {
        xin[k] = x + k >= 1 && x + k <= nx - 2;
        own[k] = XT ? x + k >= olo && x + k <= ohi : xin[k];
    }
    // WIN: rows are read through a window of planes zwin .. zwe - 1: the buffer resource starts at
    // plane zwin and covers the window only, so offsets stay 32-bit for any field size (the host
    // keeps a chunk's window, chunk_len + 6 planes, below 2 GiB).  Planes outside the window read
    // zeros.  Without WIN the window is the whole field.
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, (int)(plane * nz * sizeof(T)), 0x00020000);
    __amdgpu_buffer_rsrc_t wrsrc = rsrc;  // WIN: the current window
    int zwin = 0, zwe = nz;
    auto window = [&](int z0, int z1) {  // planes z0 .. z1, clamped to the field (wave-uniform)
        if constexpr (WIN) {
            z0 = max(z0, 0);
            z1 = min(z1, nz - 1);
            zwin = z0;
            zwe = z1 + 1;
            wrsrc = __builtin_amdgcn_make_buffer_rsrc((void *)(src + (size_t)z0 * plane), (short)0,
                                                      (int)((size_t)(zwe - zwin) * plane * sizeof(T)), 0x00020000);
        }
    };
    constexpr unsigned kOut = 0x80000000u;  // beyond every window served here: the load returns zeros

    T col[3][2][VX];  // level-0 column rows (own, tile halo) of planes p-1, p, p+1
    T rc[2][2][VX];   // ring: column rows of the two planes in flight
    // side rows: plane p+1 issued at step p, or for periodic FP64 a two-slot ring like the column
    // rows (plane p+2 issued at step p; 512^3: periodic FP64 557 -> 567 GLUP/s, but Dirichlet FP64
    // and FP32 measured 0.6-1.5 % slower with it)
    constexpr bool kSideRing = Tl::HOIST && PER && sizeof(T) == 8;
    T sid[2][VX];                        // side rows of plane p
    T rsd[kSideRing ? 2 : 1][2][VX];  // side rows in flight
    T l1o[3][VX];
    T l1h[2][VX];
    T hx[2] = {T(0), T(0)}, hxn[2] = {T(0), T(0)};  // XT: level-0 values beside the tile (rows y, yc1)

    // a lane's column as a vector offset, the row as a scalar one; outside the field the vector
    // offset carries kOut and the scalar one is 0
    const unsigned vo_in = (unsigned)x * (unsigned)sizeof(T), vo_out = vo_in | kOut;
    auto ldrow = [&](int z, int yy, bool use, T (&v)[VX]) {
        if constexpr (WIN) {
            const bool ok = use && z >= zwin && z < zwe && yy >= 0 && yy < ny;  // wave-uniform
            const unsigned so = ok ? (unsigned)(((unsigned)(z - zwin) * (unsigned)ny + (unsigned)yy) * (unsigned)nx) *
                                         (unsigned)sizeof(T)
                                   : 0u;
            const unsigned vo = ok ? vo_in : vo_out;
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                const auto t = __builtin_bit_cast(
                    typename Vec16<T>::type, __builtin_amdgcn_raw_buffer_load_b128(wrsrc, (int)(vo + 16 * q), (int)so, 0));
#pragma unroll
                for (int e = 0; e < E; ++e) v[q * E + e] = t[e];
            }
        } else {
            const bool ok = use && z >= 0 && z < nz && yy >= 0 && yy < ny;  // wave-uniform
            const unsigned so = ok ? (unsigned)(((unsigned)z * (unsigned)ny + (unsigned)yy) * (unsigned)nx) *
                                         (unsigned)sizeof(T)
                                   : 0u;
            const unsigned vo = ok ? vo_in : vo_out;
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                const auto t = __builtin_bit_cast(
                    typename Vec16<T>::type, __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)(vo + 16 * q), (int)so, 0));
#pragma unroll
                for (int e = 0; e < E; ++e) v[q * E + e] = t[e];
            }
        }
    };
    auto ld_col = [&](int z, T (&v)[2][VX]) {
        ldrow(z, y, true, v[0]);
        ldrow(z, yc1, has_h, v[1]);
    };
    auto ld_side = [&](int z, T (&v)[2][VX]) {
        ldrow(z, ys0, true, v[0]);
        ldrow(z, ys1, true, v[1]);
    };
    // XT: column X0 - 1 for lane 0 of the west wave, X0 + TW for lane 63 of the east wave (zero
    // where that column lies outside the field: there the edge lane holds a boundary column,
    // whose value is kept)
    const unsigned hvo = (XT && wx == 0 && lane == 0 && X0 >= 1)                ? (unsigned)(X0 - 1) * sizeof(T)
                         : (XT && wx == WX - 1 && lane == 63 && X0 + Tl::TW < nx) ? (unsigned)(X0 + Tl::TW) * sizeof(T)
                                                                                   : kOut;
    auto ld_halo = [&](int z, T (&h)[2]) {
        if constexpr (XT) {
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int yy = sl == 0 ? y : yc1;
                static_assert(!XT || WIN, "x tiles read through plane windows");
                const bool ok = (sl == 0 || has_h) && z >= zwin && z < zwe && yy >= 0 && yy < ny;  // wave-uniform
                const unsigned so = ok ? (unsigned)(((unsigned)(z - zwin) * (unsigned)ny + (unsigned)yy) *
                                                    (unsigned)nx) *
                                             (unsigned)sizeof(T)
                                       : 0u;
                const unsigned vo = ok ? hvo : kOut;
                if constexpr (sizeof(T) == 8)
                    h[sl] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(wrsrc, (int)vo, (int)so, 0));
                else
                    h[sl] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(wrsrc, (int)vo, (int)so, 0));
            }
        }
    };
    // x edges of the level-0 rows of plane p+1: every lane stores one value into its own slot
    // (lane 0 its first column, the others their last; no masked stores), the neighbouring wave
    // reads the slot of lane 63 or lane 0 (a broadcast read)
    auto put_l0e = [&](typename Tl::Buf &B, const T (&v)[2][VX]) {
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) B.l0x[wy][wx][sl][lane] = lane_w ? v[sl][0] : v[sl][VX - 1];
    };
    // west / east neighbours of the lane's first and last column: DPP inside the wave, the
    // neighbouring wave's edge at lanes 0 / 63.  At the tile's own x ends (one tile spans the
    // rows) those lanes hold column 0 or columns >= nx-1, whose outputs keep their stored value,
    // so whatever the clamped read there returns is never used.  Both candidates are computed
    // first (a DPP move inside one arm of the select would make it a divergent branch).
    auto xnb = [&](const T (&cr)[VX], T wn, T en, T &west, T &east) {
        const T dw = lane_up1z(cr[VX - 1]);
        const T de = lane_dn1z(cr[0]);
        west = lane_w ? wn : dw;
        east = lane_e ? en : de;
    };
    auto l1row = [&](const typename Tl::Buf &B, int sl, int yy, bool pin, const T (&cr)[VX], const T (&N)[VX],
                     const T (&S)[VX], const T (&U)[VX], const T (&D)[VX], T (&out)[VX], bool norm) {
        T west, east, wn = B.l0x[wy][wx > 0 ? wx - 1 : 0][sl][63], en = B.l0x[wy][wx < WX - 1 ? wx + 1 : 0][sl][0];
        if constexpr (XT) {  // the tile's own x ends take the loaded neighbours (wave-uniform)
            if (wx == 0) wn = hx[sl];
            if (wx == WX - 1) en = hx[sl];
        }
        xnb(cr, wn, en, west, east);
        if (pin && yy >= 1 && yy <= ny - 2) {  // wave-uniform
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                const T Ev = (k < VX - 1) ? cr[k + 1] : east;
                const T Wv = (k > 0) ? cr[k - 1] : west;
                const T v = c * (Ev + Wv + N[k] + S[k] + U[k] + D[k]);
                out[k] = (FULL && k > 0 && k < VX - 1) ? v : (xin[k] ? v : cr[k]);
                if (norm) {  // wave-uniform; adding +0.0 outside the owned columns leaves the sum's bits alone
                    const double d = (double)v - (double)cr[k];
                    n0 += own[k] ? d * d : 0.0;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) out[k] = cr[k];
        }
    };
    auto level1 = [&](const typename Tl::Buf &Bp, bool pin, bool nrm) {
        if (top) {
            l1row(Bp, 0, y, pin, col[1][0], sid[1], col[1][1], col[2][0], col[0][0], l1o[2], nrm);
            l1row(Bp, 1, y - 1, pin, col[1][1], col[1][0], sid[0], col[2][1], col[0][1], l1h[1], false);
        } else if (bot) {
            l1row(Bp, 0, y, pin, col[1][0], col[1][1], sid[0], col[2][0], col[0][0], l1o[2], nrm);
            l1row(Bp, 1, y + 1, pin, col[1][1], sid[1], col[1][0], col[2][1], col[0][1], l1h[1], false);
        } else {
            l1row(Bp, 0, y, pin, col[1][0], sid[1], sid[0], col[2][0], col[0][0], l1o[2], nrm);
        }
    };
    // plain stores: the lane's 16-byte pieces are merged in the L2 (non-temporal ones were
    // counted at 1.126 x the written field, and measured no faster)
    auto store_row = [&](int z, const T (&out)[VX]) {
        T *drow = dst + (size_t)z * plane + (size_t)y * nx;
        if (FULL || (x >= olo && x + VX - 1 <= ohi)) {
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                typename Vec16<T>::type t;
#pragma unroll
                for (int e = 0; e < E; ++e) t[e] = out[q * E + e];
                *reinterpret_cast<typename Vec16<T>::type *>(drow + x + q * E) = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k)
                if (own[k]) drow[x + k] = out[k];
        }
    };
    auto store_copy = [&](int z, const T (&out)[VX]) {
        if (y >= ny) return;  // wave-uniform
        T *drow = dst + (size_t)z * plane + (size_t)y * nx;
        if (FULL) {
#pragma unroll
            for (int q = 0; q < VX / E; ++q) {
                typename Vec16<T>::type t;
#pragma unroll
                for (int e = 0; e < E; ++e) t[e] = out[q * E + e];
                *reinterpret_cast<typename Vec16<T>::type *>(drow + x + q * E) = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k)
                if (x + k >= clo && x + k <= chi) drow[x + k] = out[k];
        }
    };

    int p0 = single ? zb : zb - 1;
    const int p1 = single ? ze - 1 : ze;  // inclusive
    if (PER && !single && zb == 1) {
        // level 1 on plane 0 = level 1 on plane nz-2, from stored planes nz-3 .. nz-1
        window(nz - 3, nz - 1);
        ld_col(nz - 3, col[0]);
        ld_col(nz - 2, col[1]);
        ld_col(nz - 1, col[2]);
        ld_side(nz - 2, sid);
        ld_halo(nz - 2, hx);
        put_l0e(lds[0], col[1]);
        __syncthreads();
        level1(lds[0], true, false);
#pragma unroll
        for (int k = 0; k < VX; ++k) {
            l1o[1][k] = l1o[2][k];
            l1h[0][k] = l1h[1][k];
        }
        p0 = 1;
    }
    window(p0 - 1, p1 + 3);  // every plane the loop reads (side rows one plane ahead, columns three)
    ld_col(p0 - 1, col[0]);
    ld_col(p0, col[1]);
    ld_col(p0 + 1, col[2]);
    ld_side(p0, sid);
    ld_halo(p0, hx);
    ld_col(p0 + 2, rc[1]);  // slot of step p0 + 1, moved into the window at the end of step p0
    if constexpr (kSideRing) ld_side(p0 + 1, rsd[1]);
    put_l0e(lds[p0 % 3], col[1]);
    __syncthreads();

    // one plane step; WRAP: the periodic last step of the chunk ending at plane nz-2, where level
    // 1 on plane nz-1 is level 1 on plane 1, from stored planes 0 .. 2 (run after the loop, so
    // the loop body carries none of it)
    auto step = [&](int p, auto slot, auto wrap) {
        constexpr int S = decltype(slot)::value;
        constexpr bool WRAP = decltype(wrap)::value;
        typename Tl::Buf &Bp = lds[p % 3];
        bool pin = p >= 1 && p <= nz - 2;
        if constexpr (WRAP) {
            window(0, 2);
            ld_col(0, col[0]);
            ld_col(1, col[1]);
            ld_col(2, col[2]);
            ld_side(1, sid);
            ld_halo(1, hx);
            put_l0e(Bp, col[1]);
            __syncthreads();
            pin = true;
        } else {
            if constexpr (kSideRing)
                ld_side(p + 2, rsd[S]);  // into the window at the end of step p + 1
            else
                ld_side(p + 1, rsd[0]);  // into the window at the end of this step (issued first)
            ld_halo(p + 1, hxn);
            ld_col(p + 3, rc[S]);  // into the window at the end of step p + 1
        }
        const bool nrm = want0 && p >= zb && p < ze;
        // level-2 inputs (two-sweep passes): the level-1 rows of plane p - 1, published at step p - 1
        // before the barrier that ended it, are read before level 1 of this step so that their LDS
        // latency overlaps its arithmetic
        const int zo = p - 1;
        const bool l2 = !single && zo >= zb && zo >= 1 && zo <= nz - 2 && y >= 1 && y <= ny - 2;  // wave-uniform
        const typename Tl::Buf &R = lds[(p + 2) % 3];
        T N[VX], Sv[VX], wn2 = T(0), en2 = T(0);
        auto l2_reads = [&] {
            if (l2) {
                if (!top) ld_vec<T, VX>(&R.row[wy - 1 < 0 ? 0 : wy - 1][xl], Sv);
                if (!bot) ld_vec<T, VX>(&R.row[wy + 1 >= WY ? WY - 1 : wy + 1][xl], N);
                wn2 = R.row[wy][wx > 0 ? wx * Tl::SEG - 1 : 0];
                en2 = R.row[wy][wx < WX - 1 ? (wx + 1) * Tl::SEG : 0];
            }
        };
        if constexpr (kHoist) l2_reads();
        level1(Bp, pin, nrm);
        typename Tl::Buf &Bn = lds[(p + 1) % 3];
        if (single) {
            if (y >= 1 && y <= ny - 2) store_row(p, l1o[2]);
            if (PER) {
                if (p == nz - 2) store_copy(0, l1o[2]);
                if (p == 1) store_copy(nz - 1, l1o[2]);
            }
            put_l0e(Bn, col[2]);
            __syncthreads();
        } else {
            st_vec<T, VX>(&Bp.row[wy][xl], l1o[2]);  // (its x edges are read from here too)
            put_l0e(Bn, col[2]);
            if constexpr (!kHoist) {
                __syncthreads();
                l2_reads();
            }
            if (l2) {
                if (top) {
#pragma unroll
                    for (int k = 0; k < VX; ++k) Sv[k] = l1h[0][k];
                }
                if (bot) {
#pragma unroll
                    for (int k = 0; k < VX; ++k) N[k] = l1h[0][k];
                }
                const T (&cr)[VX] = l1o[1];
                T west, east;
                xnb(cr, wn2, en2, west, east);
                T out[VX];
#pragma unroll
                for (int k = 0; k < VX; ++k) {
                    const T Ev = (k < VX - 1) ? cr[k + 1] : east;
                    const T Wv = (k > 0) ? cr[k - 1] : west;
                    const T v = c * (Ev + Wv + N[k] + Sv[k] + l1o[2][k] + l1o[0][k]);
                    out[k] = (FULL && k > 0 && k < VX - 1) ? v : (xin[k] ? v : cr[k]);
                    if (want1) {
                        const double d = (double)v - (double)cr[k];
                        n1 += own[k] ? d * d : 0.0;
                    }
                }
                store_row(zo, out);
                if (PER) {
                    if (zo == nz - 2) store_copy(0, out);
                    if (zo == 1) store_copy(nz - 1, out);
                }
            } else if (PER && zo >= zb && (zo == 1 || zo == nz - 2) && (y == 0 || y == ny - 1)) {
                if (zo == nz - 2) store_copy(0, l1o[1]);
                if (zo == 1) store_copy(nz - 1, l1o[1]);
            }
            if constexpr (kHoist) __syncthreads();  // this step's rows and edges are published for the next step
        }
        if constexpr (WRAP) return;  // the chunk's last step
        // ---- slide: the window takes the ring slot issued one step earlier
#pragma unroll
        for (int k = 0; k < VX; ++k) {
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                col[0][sl][k] = col[1][sl][k];
                col[1][sl][k] = col[2][sl][k];
                col[2][sl][k] = rc[S ^ 1][sl][k];
                sid[sl][k] = rsd[kSideRing ? (S ^ 1) : 0][sl][k];
            }
            l1o[0][k] = l1o[1][k];
            l1o[1][k] = l1o[2][k];
            l1h[0][k] = l1h[1][k];
        }
        if constexpr (XT) {
            hx[0] = hxn[0];
            hx[1] = hxn[1];
        }
    };
    if constexpr (PER) {
        const bool wrap_top = !single && ze == nz - 1;
        const int pl = wrap_top ? p1 - 1 : p1;
        for (int p = p0; p <= pl;) {
            step(p, Slot<0>{}, std::false_type{});
            if (++p > pl) break;
            step(p, Slot<1>{}, std::false_type{});
            if (++p > pl) break;
        }
        if (wrap_top) step(p1, Slot<0>{}, std::true_type{});
    } else {
        // (this loop form measured 2 % faster for the Dirichlet build than the one above)
        for (int p = p0;;) {
            step(p, Slot<0>{}, std::false_type{});
            if (++p > p1) break;
            step(p, Slot<1>{}, std::false_type{});
            if (++p > p1) break;
        }
    }
    __syncthreads();  // the LDS buffers are reused by the next unit
}


template <typename T, int WX, int WY, int VXT, bool ALIGNED, bool FULL, bool PER, bool TIMED, bool XT = false,
          bool WIN = false>
__global__ void __launch_bounds__(64 * WX * WY) jacobi3d_t2_kernel(JacobiArgs A) {
    static_assert(!(XT || WIN) || ALIGNED, "x tiles and plane windows: the buffer-load build");
    using Tl = T2Tile<T, WX, WY, VXT>;
    __shared__ typename Tl::Buf bufs[3];
    __shared__ double red[Tl::BLOCK / 64];
    const int g = gridDim.x, b = blockIdx.x;
    const int lb = (g & 7) == 0 ? (b & 7) * (g >> 3) + (b >> 3) : b;  // XCD-contiguous order
    const int nunits = A.ntx * A.nty * A.chunks;  // (ntx = 1 unless XT)
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    int pass = 0;
    for (int k = 0; k < A.niter; ++pass) {
        const int steps = min(2, A.niter - k);
        const int it = A.it0 + k;
        const bool want0 = A.norm_every > 0 && (it % A.norm_every) == 0;
        const bool want1 = steps == 2 && A.norm_every > 0 && ((it + 1) % A.norm_every) == 0;
        const T *src = static_cast<const T *>((pass & 1) ? A.buf1 : A.buf0);
        T *dst = static_cast<T *>((pass & 1) ? A.buf0 : A.buf1);
        double n0 = 0.0, n1 = 0.0;
        for (int u = lb; u < nunits; u += g) {
            const int zc = u / (A.nty * A.ntx);
            const int r = u - zc * (A.nty * A.ntx);
            const int ty = r / A.ntx, tx = r - ty * A.ntx;
            // 16-byte aligned rows: the two-step lookahead through buffer loads (fields below 2 GiB,
            // which the host guarantees for the ALIGNED builds)
            if constexpr (ALIGNED)
                sweep3d_t2r<T, WX, WY, VXT, FULL, PER, XT, WIN>(src, dst, A, tx, ty, zc, steps == 1, want0, want1,
                                                                n0, n1, bufs);
            else
                sweep3d_t2<T, WX, WY, VXT, ALIGNED, FULL, PER>(src, dst, A, ty, zc, steps == 1, want0, want1, n0,
                                                               n1, bufs);
        }
        if (want0) {
            const double s = block_sum_waves<Tl::BLOCK / 64>(n0, red);
            if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
        }
        if (want1) {
            const double s = block_sum_waves<Tl::BLOCK / 64>(n1, red);
            if (threadIdx.x == 0) A.norm_part[(size_t)((it + 1) / A.norm_every) * gridDim.x + blockIdx.x] = s;
        }
        k += steps;
        if (k < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
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
#pragma unroll
                    for (int e = 0; e < VX; ++e) t[e] = out[e];
                    store16(reinterpret_cast<typename Vec16<T>::type *>(drow + x), t);
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

template <typename T, int RB, bool ALIGNED, int BLOCK, bool TIMED>
__global__ void __launch_bounds__(BLOCK) jacobi2d_kernel(JacobiArgs A) {
    __shared__ double lds[BLOCK / 64];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int nunits = A.ntx * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
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
                const double s = block_sum_waves<BLOCK / 64>(nacc, lds);
                if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        if (k + 1 < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// ------------------------------------------------- 2D, two sweeps per pass --
// Temporal blocking for the 2D Dirichlet case (the reference's PERKS 2D
// caches the slab across iterations, Stencil/jacobi2D/PERKS/
// jacobi-general-kernel.cuh:22-505).  Each wavefront owns SEG = 64*VX columns
// and streams a chunk of rows: at step r it computes sweep t+1 ("level 1") on
// row r from level-0 rows r-1, r, r+1 and sweep t+2 on row r-1 from level-1
// rows r-2, r-1, r, then stores only the latter.  Wavefronts are independent:
// lane 0 and lane 63 also compute level 1 on the column just outside the
// segment (from two extra level-0 columns per row), and every chunk
// recomputes level 1 on the row just outside each end.  Same expressions as
// one sweep, boundary points kept: bit-identical to two single sweeps.
// columns per lane of the 2D two-sweep kernel: one 16-byte vector in
// 1024-thread workgroups, two in 512-thread workgroups (room for the registers)
template <typename T>
constexpr int t2d_cols(int block) {
    return (block == 512 ? 2 : 1) * 16 / (int)sizeof(T);
}

// One 16-byte vector per lane leaves as a non-temporal store.  With two or
// more per lane each store instruction covers every other 16 bytes of the row,
// and non-temporal pieces of one 64-byte block are not always merged before
// they leave the cache; for FP64 plain stores let the L2 merge them (16384^2,
// four sweeps per pass: counted writes 1.051 -> 1.000 x the field, 975 -> 983
// GLUP/s).  FP32 rows measured faster non-temporal (1614 vs 1606 GLUP/s).
template <typename T, int VX, bool ALIGNED>
__device__ __forceinline__ void store_row_nt(T *drow, int x, int nx, const T (&v)[VX], const bool (&xin)[VX]) {
    constexpr int E = 16 / sizeof(T);
    if (ALIGNED && x >= 1 && x + VX - 1 <= nx - 2) {
#pragma unroll
        for (int q = 0; q < VX / E; ++q) {
            typename Vec16<T>::type t;
#pragma unroll
            for (int e = 0; e < E; ++e) t[e] = v[q * E + e];
            auto *dp = reinterpret_cast<typename Vec16<T>::type *>(drow + x + q * E);
            if constexpr (sizeof(T) == 8 && VX / E > 1)
                *dp = t;
            else
                store16(dp, t);
        }
    } else {
#pragma unroll
        for (int k = 0; k < VX; ++k)
            if (xin[k]) drow[x + k] = v[k];
    }
}

// PER: the reference's periodic y (after every sweep row ny-2 is copied to row 0 and row 1 to
// row ny-1), as in the 3D two-sweep kernel: level 1 on row 0 is level 1 on row ny-2 (the chunk
// starting at row 1 computes it first, from stored rows ny-3 .. ny-1), level 1 on row ny-1 is
// level 1 on row 1 (the chunk ending at row ny-2 computes it in its last step, from stored rows
// 0 .. 2), and whoever stores rows ny-2 and 1 also writes them whole into rows 0 and ny-1.
template <typename T, int VX, bool ALIGNED, int LA = 2, bool PER = false>
__device__ __forceinline__ void sweep2d_t2(const T *__restrict__ src, T *__restrict__ dst, const JacobiArgs &A,
                                           int tx, int yc, bool single, bool want0, bool want1, double &n0,
                                           double &n1) {
    constexpr int SEG = 64 * VX;
    constexpr int NW = 3 + LA;  // level-0 rows r-1 .. r+1+LA (the last LA in flight)
    const int lane = threadIdx.x & 63;
    const int nx = A.nx, ny = A.ny;
    const int x0 = tx * SEG;
    const int x = x0 + lane * VX;
    const int yb = 1 + yc * A.chunk_len;
    const int ye = min(yb + A.chunk_len, ny - 1);
    if (yb >= ye) return;  // uniform over the wavefront
    const T quarter = T(0.25);
    const bool lane_w = lane == 0, lane_e = lane == 63;
    // edge columns of this lane: lane 0 -> x0-2, x0-1; lane 63 -> x0+SEG, x0+SEG+1
    const int ex1 = lane_w ? x0 - 1 : x0 + SEG;      // adjacent column (level-1 edge column)
    const int ex2 = lane_w ? x0 - 2 : x0 + SEG + 1;  // the one beyond
    const bool edge_lane = lane_w || lane_e;
    bool xin[VX];
#pragma unroll
    for (int k = 0; k < VX; ++k) xin[k] = x + k >= 1 && x + k <= nx - 2;
    const bool ein = edge_lane && ex1 >= 1 && ex1 <= nx - 2;  // the level-1 edge column is interior

    T a[NW][VX];       // level-0 rows
    T e1[NW], e2[NW];  // level-0 edge columns ex1, ex2 of those rows (edge lanes)
    T l[3][VX] = {};  // level-1 rows r-2, r-1, r
    T le[3] = {};     // level-1 edge column ex1 of those rows (edge lanes)

    auto load = [&](int y, T (&v)[VX], T &v1, T &v2) {
        const bool ok = y >= 0 && y < ny;  // uniform
        const T *pr = src + (size_t)(ok ? y : 0) * nx;
        load_row<T, VX, ALIGNED>(pr, x, nx, ok, v);
        v1 = (ok && edge_lane && ex1 >= 0 && ex1 < nx) ? pr[ex1] : T(0);
        v2 = (ok && edge_lane && ex2 >= 0 && ex2 < nx) ? pr[ex2] : T(0);
    };

    // level 1 on one row into l[2] / le[2], from the level-0 rows above (u), at (c) and below (d)
    auto level1 = [&](const T (&u)[VX], const T (&cr)[VX], const T (&d)[VX], T eu, T ec1, T ec2, T ed, bool rin,
                      bool nrm) {
        T west = lane_up1(cr[VX - 1]);
        T east = lane_dn1(cr[0]);
        if (lane_w) west = ec1;
        if (lane_e) east = ec1;
        if (rin) {  // wave-uniform
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                const T E = (k < VX - 1) ? cr[k + 1] : east;
                const T W = (k > 0) ? cr[k - 1] : west;
                const T v = quarter * (E + W + d[k] + u[k]);
                l[2][k] = xin[k] ? v : cr[k];
                if (nrm && xin[k]) {
                    const double dd = (double)v - (double)cr[k];
                    n0 += dd * dd;
                }
            }
            // edge column: lane 0 -> x0-1 (east neighbour x0 = cr[0]); lane 63 -> x0+SEG (west x0+SEG-1)
            const T E = lane_w ? cr[0] : ec2;
            const T W = lane_w ? ec2 : cr[VX - 1];
            const T v = quarter * (E + W + ed + eu);
            le[2] = ein ? v : ec1;
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) l[2][k] = cr[k];
            le[2] = ec1;
        }
    };
    // PER: a row stored whole (boundary columns too) as the halo row's copy
    auto store_copy = [&](int row, const T (&v)[VX]) {
        T *drow = dst + (size_t)row * nx;
#pragma unroll
        for (int k = 0; k < VX; ++k)
            if (x + k < nx) drow[x + k] = v[k];
    };
    // PER: level 1 of a wrapped halo row from three stored rows, read into the first three
    // window slots (before the window is filled, or in the chunk's last step, after which they are
    // not read); no norm: its row is another chunk's
    auto level1_from = [&](int ru) {
#pragma unroll
        for (int i = 0; i < 3; ++i) load(ru + i, a[i], e1[i], e2[i]);
        level1(a[0], a[1], a[2], e1[0], e1[1], e2[1], e1[2], true, false);
    };

    int r0 = single ? yb : yb - 1;
    const int r1 = single ? ye - 1 : ye;  // inclusive
    if (PER && !single && yb == 1) {
        // level 1 on row 0 = level 1 on row ny-2, from stored rows ny-3 .. ny-1
        level1_from(ny - 3);
#pragma unroll
        for (int k = 0; k < VX; ++k) l[1][k] = l[2][k];
        le[1] = le[2];
        r0 = 1;
    }
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) load(r0 - 1 + i, a[i], e1[i], e2[i]);
    // after level 1 on row r: level 2 on row r-1 (or the single sweep's store), then the slide
    auto finish = [&](int r) {
        const bool rin = r >= 1 && r <= ny - 2;
        if (single) {
            if (rin) store_row_nt<T, VX, ALIGNED>(dst + (size_t)r * nx, x, nx, l[2], xin);
            if (PER) {
                if (r == ny - 2) store_copy(0, l[2]);
                if (r == 1) store_copy(ny - 1, l[2]);
            }
        } else if (r - 1 >= yb && r - 1 >= 1 && r - 1 <= ny - 2) {
            // ---- level 2 on row r-1
            const T *cr = l[1];
            T west = lane_up1(cr[VX - 1]);
            T east = lane_dn1(cr[0]);
            if (lane_w) west = le[1];
            if (lane_e) east = le[1];
            T out[VX];
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                const T E = (k < VX - 1) ? cr[k + 1] : east;
                const T W = (k > 0) ? cr[k - 1] : west;
                const T v = quarter * (E + W + l[2][k] + l[0][k]);
                out[k] = xin[k] ? v : cr[k];
                if (want1 && xin[k]) {
                    const double d = (double)v - (double)cr[k];
                    n1 += d * d;
                }
            }
            store_row_nt<T, VX, ALIGNED>(dst + (size_t)(r - 1) * nx, x, nx, out, xin);
            if (PER) {
                if (r - 1 == ny - 2) store_copy(0, out);
                if (r - 1 == 1) store_copy(ny - 1, out);
            }
        }
        // ---- slide
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) {
#pragma unroll
            for (int k = 0; k < VX; ++k) a[i][k] = a[i + 1][k];
            e1[i] = e1[i + 1];
            e2[i] = e2[i + 1];
        }
#pragma unroll
        for (int k = 0; k < VX; ++k) {
            l[0][k] = l[1][k];
            l[1][k] = l[2][k];
        }
        le[0] = le[1];
        le[1] = le[2];
    };
    // PER: the chunk ending at row ny-2 takes its last step (level 1 on row ny-1 = level 1 on row
    // 1, from stored rows 0 .. 2) after the loop, so the loop body carries none of it
    const bool wrap = PER && !single && r1 == ny - 1;
    const int rl = wrap ? r1 - 1 : r1;
    for (int r = r0; r <= rl; ++r) {
        load(r + 1 + LA, a[NW - 1], e1[NW - 1], e2[NW - 1]);  // consumed LA steps later
        // ---- level 1 on row r (own columns, and the edge column for lanes 0 / 63)
        level1(a[0], a[1], a[2], e1[0], e1[1], e2[1], e1[2], r >= 1 && r <= ny - 2, want0 && r >= yb && r < ye);
        finish(r);
    }
    if (wrap) {
        level1_from(0);
        finish(r1);
    }
}

template <typename T, bool ALIGNED, int BLOCK, bool TIMED, bool PER = false>
__global__ void __launch_bounds__(BLOCK) jacobi2d_t2_kernel(JacobiArgs A) {
    constexpr int VX = t2d_cols<T>(BLOCK);
    __shared__ double lds[BLOCK / 64];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int nunits = A.ntx * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    int pass = 0;
    for (int k = 0; k < A.niter; ++pass) {
        const int steps = min(2, A.niter - k);
        const int it = A.it0 + k;
        const bool want0 = A.norm_every > 0 && (it % A.norm_every) == 0;
        const bool want1 = steps == 2 && A.norm_every > 0 && ((it + 1) % A.norm_every) == 0;
        const T *src = static_cast<const T *>((pass & 1) ? A.buf1 : A.buf0);
        T *dst = static_cast<T *>((pass & 1) ? A.buf0 : A.buf1);
        double n0 = 0.0, n1 = 0.0;
        for (int u = wave_id; u < nunits; u += nwaves) {
            const int yc = u / A.ntx;
            const int tx = u - yc * A.ntx;
            sweep2d_t2<T, VX, ALIGNED, 2, PER>(src, dst, A, tx, yc, steps == 1, want0, want1, n0, n1);
        }
        if (want0) {
            const double s = block_sum_waves<BLOCK / 64>(n0, lds);
            if (threadIdx.x == 0) A.norm_part[(size_t)(it / A.norm_every) * gridDim.x + blockIdx.x] = s;
        }
        if (want1) {
            const double s = block_sum_waves<BLOCK / 64>(n1, lds);
            if (threadIdx.x == 0) A.norm_part[(size_t)((it + 1) / A.norm_every) * gridDim.x + blockIdx.x] = s;
        }
        k += steps;
        if (k < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// Three sweeps per pass in 2D: as sweep2d_t2 with one more level.  At step r
// the wavefront computes level 1 on row r, level 2 on row r-1 and level 3 on
// row r-2 (stored).  Lanes 0 and 63 carry three level-0 columns outside the
// segment (e1 adjacent, e2, e3), level 1 on e1 and e2 and level 2 on e1; each
// chunk recomputes levels 1-2 on the two rows beyond each end.  Every value
// comes from the single-sweep expression: bit-identical to three sweeps.
template <typename T, int VX, bool ALIGNED, int LA = 2>
__device__ __forceinline__ void sweep2d_t3(const T *__restrict__ src, T *__restrict__ dst, const JacobiArgs &A,
                                           int tx, int yc, bool want0, bool want1, bool want2, double &n0,
                                           double &n1, double &n2) {
    constexpr int SEG = 64 * VX;
    constexpr int NW = 3 + LA;  // level-0 rows r-1 .. r+1+LA (the last LA in flight)
    const int lane = threadIdx.x & 63;
    const int nx = A.nx, ny = A.ny;
    const int x0 = tx * SEG;
    const int x = x0 + lane * VX;
    const int yb = 1 + yc * A.chunk_len;
    const int ye = min(yb + A.chunk_len, ny - 1);
    if (yb >= ye) return;  // uniform over the wavefront
    const T quarter = T(0.25);
    const bool lane_w = lane == 0, lane_e = lane == 63;
    const bool edge_lane = lane_w || lane_e;
    // edge columns: lane 0 -> x0-1, x0-2, x0-3; lane 63 -> x0+SEG, +1, +2
    int ec[3];
    bool ein[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ec[i] = lane_w ? x0 - 1 - i : x0 + SEG + i;
        ein[i] = edge_lane && ec[i] >= 1 && ec[i] <= nx - 2;
    }
    bool xin[VX];
#pragma unroll
    for (int k = 0; k < VX; ++k) xin[k] = x + k >= 1 && x + k <= nx - 2;

    T a[NW][VX];     // level-0 rows
    T ae[NW][3];     // their edge columns e1, e2, e3 (edge lanes)
    T l1[3][VX] = {}, l1e1[3] = {}, l1e2[3] = {};  // level 1, rows r-2, r-1, r
    T l2[3][VX] = {}, l2e1[3] = {};                // level 2, rows r-3, r-2, r-1

    auto load = [&](int y, T (&v)[VX], T (&e)[3]) {
        const bool ok = y >= 0 && y < ny;  // uniform
        const T *pr = src + (size_t)(ok ? y : 0) * nx;
        load_row<T, VX, ALIGNED>(pr, x, nx, ok, v);
#pragma unroll
        for (int i = 0; i < 3; ++i) e[i] = (ok && edge_lane && ec[i] >= 0 && ec[i] < nx) ? pr[ec[i]] : T(0);
    };
    // one row of a level from the row below/above (S/N) and its lanes' edge value
    auto row = [&](const T (&cr)[VX], const T (&N)[VX], const T (&S)[VX], T wedge, T eedge, T (&out)[VX],
                   bool norm, double &acc) {
        T west = lane_up1(cr[VX - 1]);
        T east = lane_dn1(cr[0]);
        if (lane_w) west = wedge;
        if (lane_e) east = eedge;
#pragma unroll
        for (int k = 0; k < VX; ++k) {
            const T E = (k < VX - 1) ? cr[k + 1] : east;
            const T W = (k > 0) ? cr[k - 1] : west;
            const T v = quarter * (E + W + N[k] + S[k]);
            out[k] = xin[k] ? v : cr[k];
            if (norm && xin[k]) {
                const double d = (double)v - (double)cr[k];
                acc += d * d;
            }
        }
    };

    const int r0 = yb - 2, r1 = ye + 1;  // level-1 rows
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) load(r0 - 1 + i, a[i], ae[i]);
    for (int r = r0; r <= r1; ++r) {
        load(r + 1 + LA, a[NW - 1], ae[NW - 1]);  // consumed LA steps later
        // ---- level 1 on row r (own columns; e1 and e2 for the edge lanes)
        T n1r[VX], n1e1, n1e2;
        if (r >= 1 && r <= ny - 2) {  // wave-uniform
            row(a[1], a[2], a[0], ae[1][0], ae[1][0], n1r, want0 && r >= yb && r < ye, n0);
            {
                const T E = lane_w ? a[1][0] : ae[1][1];
                const T W = lane_w ? ae[1][1] : a[1][VX - 1];
                const T v = quarter * (E + W + ae[2][0] + ae[0][0]);
                n1e1 = ein[0] ? v : ae[1][0];
            }
            {
                const T E = lane_w ? ae[1][0] : ae[1][2];
                const T W = lane_w ? ae[1][2] : ae[1][0];
                const T v = quarter * (E + W + ae[2][1] + ae[0][1]);
                n1e2 = ein[1] ? v : ae[1][1];
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) n1r[k] = a[1][k];
            n1e1 = ae[1][0];
            n1e2 = ae[1][1];
        }
#pragma unroll
        for (int k = 0; k < VX; ++k) {
            l1[0][k] = l1[1][k];
            l1[1][k] = l1[2][k];
            l1[2][k] = n1r[k];
        }
        l1e1[0] = l1e1[1];
        l1e1[1] = l1e1[2];
        l1e1[2] = n1e1;
        l1e2[0] = l1e2[1];
        l1e2[1] = l1e2[2];
        l1e2[2] = n1e2;
        // ---- level 2 on row r-1 (own columns; e1 for the edge lanes)
        const int q = r - 1;
        if (q >= yb - 1 && q <= ye) {  // wave-uniform
            T n2r[VX], n2e1;
            if (q >= 1 && q <= ny - 2) {
                row(l1[1], l1[2], l1[0], l1e1[1], l1e1[1], n2r, want1 && q >= yb && q < ye, n1);
                const T E = lane_w ? l1[1][0] : l1e2[1];
                const T W = lane_w ? l1e2[1] : l1[1][VX - 1];
                const T v = quarter * (E + W + l1e1[2] + l1e1[0]);
                n2e1 = ein[0] ? v : l1e1[1];
            } else {
#pragma unroll
                for (int k = 0; k < VX; ++k) n2r[k] = l1[1][k];
                n2e1 = l1e1[1];
            }
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                l2[0][k] = l2[1][k];
                l2[1][k] = l2[2][k];
                l2[2][k] = n2r[k];
            }
            l2e1[0] = l2e1[1];
            l2e1[1] = l2e1[2];
            l2e1[2] = n2e1;
        }
        // ---- level 3 on row r-2, stored
        const int q3 = r - 2;
        if (q3 >= yb && q3 < ye) {  // wave-uniform; interior rows only
            T out[VX];
            row(l2[1], l2[2], l2[0], l2e1[1], l2e1[1], out, want2, n2);
            store_row_nt<T, VX, ALIGNED>(dst + (size_t)q3 * nx, x, nx, out, xin);
        }
        // ---- slide the level-0 window
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) {
#pragma unroll
            for (int k = 0; k < VX; ++k) a[i][k] = a[i + 1][k];
#pragma unroll
            for (int e = 0; e < 3; ++e) ae[i][e] = ae[i + 1][e];
        }
    }
}

template <typename T, bool ALIGNED, int BLOCK, bool TIMED>
__global__ void __launch_bounds__(BLOCK) jacobi2d_t3_kernel(JacobiArgs A) {
    constexpr int VX = t2d_cols<T>(BLOCK);
    __shared__ double lds[BLOCK / 64];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int nunits = A.ntx * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    int pass = 0;
    for (int k = 0; k < A.niter; ++pass) {
        const int steps = min(3, A.niter - k);
        const int it = A.it0 + k;
        bool want[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) want[l] = l < steps && A.norm_every > 0 && ((it + l) % A.norm_every) == 0;
        const T *src = static_cast<const T *>((pass & 1) ? A.buf1 : A.buf0);
        T *dst = static_cast<T *>((pass & 1) ? A.buf0 : A.buf1);
        double nrm[3] = {0.0, 0.0, 0.0};
        for (int u = wave_id; u < nunits; u += nwaves) {
            const int yc = u / A.ntx;
            const int tx = u - yc * A.ntx;
            if (steps == 3)
                sweep2d_t3<T, VX, ALIGNED>(src, dst, A, tx, yc, want[0], want[1], want[2], nrm[0], nrm[1], nrm[2]);
            else
                sweep2d_t2<T, VX, ALIGNED>(src, dst, A, tx, yc, steps == 1, want[0], want[1], nrm[0], nrm[1]);
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            if (want[l]) {  // uniform over the grid
                const double s = block_sum_waves<BLOCK / 64>(nrm[l], lds);
                if (threadIdx.x == 0) A.norm_part[(size_t)((it + l) / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        k += steps;
        if (k < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// L sweeps per pass in 2D, the general form of sweep2d_t3: at step r level j
// (1..L) is computed on row r-(j-1); lanes 0 and 63 carry L level-0 columns
// beyond the segment and level j on L-j of them; each chunk recomputes the
// rows a level needs beyond its ends.  Levels are compile-time (template
// recursion), so every window stays in registers.
template <typename T, int VX, bool ALIGNED, int L, int LA = 2>
struct Sweep2dN {
    static constexpr int NW = 3 + LA;
    const T *__restrict__ src;
    T *__restrict__ dst;
    int nx, ny, yb, ye, x, lane;
    bool lane_w, lane_e, edge_lane;
    int ec[L];
    bool ein[L], xin[VX];
    T a[NW][VX], ae[NW][L];
    T lv[L - 1][3][VX], lve[L - 1][3][L];
    const bool *want;
    double *nrm;

    __device__ __forceinline__ void load(int y, T (&v)[VX], T (&e)[L]) {
        const bool ok = y >= 0 && y < ny;  // uniform
        const T *pr = src + (size_t)(ok ? y : 0) * nx;
        load_row<T, VX, ALIGNED>(pr, x, nx, ok, v);
#pragma unroll
        for (int i = 0; i < L; ++i) e[i] = (ok && edge_lane && ec[i] >= 0 && ec[i] < nx) ? pr[ec[i]] : T(0);
    }
    // one level-J row from the level below: S/C/N rows and their edge columns
    template <int J>
    __device__ __forceinline__ void level_row(const T (&S)[VX], const T (&C)[VX], const T (&N)[VX], const T (&Se)[L],
                                              const T (&Ce)[L], const T (&Ne)[L], bool rin, bool norm,
                                              T (&out)[VX], T (&oe)[L]) {
        constexpr int NE = L - J;  // edge columns computed at this level
        const T quarter = T(0.25);
        if (rin) {
            T west = lane_up1(C[VX - 1]);
            T east = lane_dn1(C[0]);
            if (lane_w) west = Ce[0];
            if (lane_e) east = Ce[0];
#pragma unroll
            for (int k = 0; k < VX; ++k) {
                const T E = (k < VX - 1) ? C[k + 1] : east;
                const T W = (k > 0) ? C[k - 1] : west;
                const T v = quarter * (E + W + N[k] + S[k]);
                out[k] = xin[k] ? v : C[k];
                if (norm && xin[k]) {
                    const double d = (double)v - (double)C[k];
                    nrm[J - 1] += d * d;
                }
            }
#pragma unroll
            for (int i = 0; i < NE; ++i) {
                const T inner = i == 0 ? (lane_w ? C[0] : C[VX - 1]) : Ce[i > 0 ? i - 1 : 0];
                const T outer = Ce[i + 1];
                const T E = lane_w ? inner : outer;
                const T W = lane_w ? outer : inner;
                const T v = quarter * (E + W + Ne[i] + Se[i]);
                oe[i] = ein[i] ? v : Ce[i];
            }
        } else {
#pragma unroll
            for (int k = 0; k < VX; ++k) out[k] = C[k];
#pragma unroll
            for (int i = 0; i < NE; ++i) oe[i] = Ce[i];
        }
    }
    template <int J>
    __device__ __forceinline__ void levels(int r) {
        if constexpr (J <= L) {
            const int q = r - (J - 1);
            if (q >= yb - (L - J) && q <= ye - 1 + (L - J)) {  // wave-uniform
                T out[VX], oe[L];
                const bool rin = q >= 1 && q <= ny - 2;
                const bool norm = want[J - 1] && q >= yb && q < ye;
                if constexpr (J == 1)
                    level_row<J>(a[0], a[1], a[2], ae[0], ae[1], ae[2], rin, norm, out, oe);
                else
                    level_row<J>(lv[J - 2][0], lv[J - 2][1], lv[J - 2][2], lve[J - 2][0], lve[J - 2][1],
                                 lve[J - 2][2], rin, norm, out, oe);
                if constexpr (J < L) {
#pragma unroll
                    for (int k = 0; k < VX; ++k) {
                        lv[J - 1][0][k] = lv[J - 1][1][k];
                        lv[J - 1][1][k] = lv[J - 1][2][k];
                        lv[J - 1][2][k] = out[k];
                    }
#pragma unroll
                    for (int i = 0; i < L - J; ++i) {
                        lve[J - 1][0][i] = lve[J - 1][1][i];
                        lve[J - 1][1][i] = lve[J - 1][2][i];
                        lve[J - 1][2][i] = oe[i];
                    }
                } else if (q >= yb && q < ye) {
                    store_row_nt<T, VX, ALIGNED>(dst + (size_t)q * nx, x, nx, out, xin);
                }
            }
            levels<J + 1>(r);
        }
    }
    __device__ __forceinline__ void run(int tx, int yc, const JacobiArgs &A) {
        constexpr int SEG = 64 * VX;
        lane = threadIdx.x & 63;
        nx = A.nx;
        ny = A.ny;
        const int x0 = tx * SEG;
        x = x0 + lane * VX;
        yb = 1 + yc * A.chunk_len;
        ye = min(yb + A.chunk_len, ny - 1);
        if (yb >= ye) return;  // uniform over the wavefront
        lane_w = lane == 0;
        lane_e = lane == 63;
        edge_lane = lane_w || lane_e;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            ec[i] = lane_w ? x0 - 1 - i : x0 + SEG + i;
            ein[i] = edge_lane && ec[i] >= 1 && ec[i] <= nx - 2;
        }
#pragma unroll
        for (int k = 0; k < VX; ++k) xin[k] = x + k >= 1 && x + k <= nx - 2;
#pragma unroll
        for (int j = 0; j < L - 1; ++j)
#pragma unroll
            for (int w = 0; w < 3; ++w) {
#pragma unroll
                for (int k = 0; k < VX; ++k) lv[j][w][k] = T(0);
#pragma unroll
                for (int i = 0; i < L; ++i) lve[j][w][i] = T(0);
            }
        const int r0 = yb - L + 1, r1 = ye + L - 2;  // level-1 rows
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) load(r0 - 1 + i, a[i], ae[i]);
        for (int r = r0; r <= r1; ++r) {
            load(r + 1 + LA, a[NW - 1], ae[NW - 1]);  // consumed LA steps later
            levels<1>(r);
#pragma unroll
            for (int i = 0; i < NW - 1; ++i) {
#pragma unroll
                for (int k = 0; k < VX; ++k) a[i][k] = a[i + 1][k];
#pragma unroll
                for (int e = 0; e < L; ++e) ae[i][e] = ae[i + 1][e];
            }
        }
    }
};

template <typename T, bool ALIGNED, int BLOCK, bool TIMED>
__global__ void __launch_bounds__(BLOCK) jacobi2d_t4_kernel(JacobiArgs A) {
    constexpr int VX = t2d_cols<T>(BLOCK);
    __shared__ double lds[BLOCK / 64];
    const int wave_id = xcd_wave_id();
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const int nunits = A.ntx * A.chunks;
    unsigned int epoch = 0;
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    int pass = 0;
    for (int k = 0; k < A.niter; ++pass) {
        const int steps = min(4, A.niter - k);
        const int it = A.it0 + k;
        bool want[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) want[l] = l < steps && A.norm_every > 0 && ((it + l) % A.norm_every) == 0;
        const T *src = static_cast<const T *>((pass & 1) ? A.buf1 : A.buf0);
        T *dst = static_cast<T *>((pass & 1) ? A.buf0 : A.buf1);
        double nrm[4] = {0.0, 0.0, 0.0, 0.0};
        for (int u = wave_id; u < nunits; u += nwaves) {
            const int yc = u / A.ntx;
            const int tx = u - yc * A.ntx;
            if (steps == 4) {
                Sweep2dN<T, VX, ALIGNED, 4> sw;
                sw.src = src;
                sw.dst = dst;
                sw.want = want;
                sw.nrm = nrm;
                sw.run(tx, yc, A);
            } else if (steps == 3) {
                sweep2d_t3<T, VX, ALIGNED>(src, dst, A, tx, yc, want[0], want[1], want[2], nrm[0], nrm[1], nrm[2]);
            } else {
                sweep2d_t2<T, VX, ALIGNED>(src, dst, A, tx, yc, steps == 1, want[0], want[1], nrm[0], nrm[1]);
            }
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            if (want[l]) {  // uniform over the grid
                const double s = block_sum_waves<BLOCK / 64>(nrm[l], lds);
                if (threadIdx.x == 0) A.norm_part[(size_t)((it + l) / A.norm_every) * gridDim.x + blockIdx.x] = s;
            }
        }
        k += steps;
        if (k < A.niter) {
            if (!pass_barrier(A, epoch, clk)) return;
        }
    }
    clk.mark(CFM_PHASE_SPMV);
    clk.flush(A.timers);
}

// ------------------------------------------------------------- dispatch ---

// four sweeps per pass: 512-thread workgroups only (the 1024-thread build spills)
template <bool TM>
static KernelFn pick_t4d_kernel(bool f64, bool aligned) {
    return f64 ? (aligned ? (KernelFn)jacobi2d_t4_kernel<double, true, 512, TM>
                          : (KernelFn)jacobi2d_t4_kernel<double, false, 512, TM>)
               : (aligned ? (KernelFn)jacobi2d_t4_kernel<float, true, 512, TM>
                          : (KernelFn)jacobi2d_t4_kernel<float, false, 512, TM>);
}

// (the FP64 1024-thread build for an unaligned row pitch needs more than the 128
// VGPRs of four waves per SIMD and would spill to scratch: that shape runs 512)
template <bool TM>
static KernelFn pick_t3d_kernel(bool f64, bool aligned, int block) {
    if (block == 512)
        return f64 ? (aligned ? (KernelFn)jacobi2d_t3_kernel<double, true, 512, TM>
                              : (KernelFn)jacobi2d_t3_kernel<double, false, 512, TM>)
                   : (aligned ? (KernelFn)jacobi2d_t3_kernel<float, true, 512, TM>
                              : (KernelFn)jacobi2d_t3_kernel<float, false, 512, TM>);
    if (f64 && !aligned) return nullptr;
    return f64 ? (KernelFn)jacobi2d_t3_kernel<double, true, 1024, TM>
               : (aligned ? (KernelFn)jacobi2d_t3_kernel<float, true, 1024, TM>
                          : (KernelFn)jacobi2d_t3_kernel<float, false, 1024, TM>);
}

template <bool TM>
static KernelFn pick_t2d_kernel(bool f64, bool aligned, int block, bool periodic) {
    if (periodic) {
        if (block == 512)
            return f64 ? (aligned ? (KernelFn)jacobi2d_t2_kernel<double, true, 512, TM, true>
                                  : (KernelFn)jacobi2d_t2_kernel<double, false, 512, TM, true>)
                       : (aligned ? (KernelFn)jacobi2d_t2_kernel<float, true, 512, TM, true>
                                  : (KernelFn)jacobi2d_t2_kernel<float, false, 512, TM, true>);
        return f64 ? (aligned ? (KernelFn)jacobi2d_t2_kernel<double, true, 1024, TM, true>
                              : (KernelFn)jacobi2d_t2_kernel<double, false, 1024, TM, true>)
                   : (aligned ? (KernelFn)jacobi2d_t2_kernel<float, true, 1024, TM, true>
                              : (KernelFn)jacobi2d_t2_kernel<float, false, 1024, TM, true>);
    }
    if (block == 512)
        return f64 ? (aligned ? (KernelFn)jacobi2d_t2_kernel<double, true, 512, TM>
                              : (KernelFn)jacobi2d_t2_kernel<double, false, 512, TM>)
                   : (aligned ? (KernelFn)jacobi2d_t2_kernel<float, true, 512, TM>
                              : (KernelFn)jacobi2d_t2_kernel<float, false, 512, TM>);
    return f64 ? (aligned ? (KernelFn)jacobi2d_t2_kernel<double, true, 1024, TM>
                          : (KernelFn)jacobi2d_t2_kernel<double, false, 1024, TM>)
               : (aligned ? (KernelFn)jacobi2d_t2_kernel<float, true, 1024, TM>
                          : (KernelFn)jacobi2d_t2_kernel<float, false, 1024, TM>);
}

// Default workgroup size per tile shape: one workgroup fills a CU (fewer
// workgroups make the device-wide barrier cheaper).  3D: WX = 4 wavefronts
// across x and block/256 wavefronts down y share halos through LDS; 8-row
// tiles run one tile per wavefront (256).  FP64 4-row tiles in 768 threads
// would exceed the VGPRs of three waves per SIMD (scratch spills): 512.
static int default_block(bool is3d, int rows, bool f64) {
    if (is3d) return rows == 1 ? 256 : rows == 2 ? 1024 : rows == 4 ? (f64 ? 512 : 768) : 256;
    return rows == 4 ? 1024 : rows == 8 ? 768 : 512;
}

template <typename T, bool AL, bool TM>
static KernelFn pick_kernel_t(bool is3d, int rows, int block) {
#define CFM_J3(R, B) (KernelFn)jacobi3d_kernel<T, R, AL, B, TM>
#define CFM_W3(R, WY) (KernelFn)jacobi3d_wg_kernel<T, R, 4, WY, AL, TM>
#define CFM_J2(R, B) (KernelFn)jacobi2d_kernel<T, R, AL, B, TM>
    if (is3d) {
        if (block == 256) {  // one independent tile per wavefront
            switch (rows) {
                case 1: return CFM_J3(1, 256);
                case 2: return CFM_J3(2, 256);
                case 4: return CFM_J3(4, 256);
                case 8: return CFM_J3(8, 256);
                default: return nullptr;
            }
        }
        // shared-tile shapes that fit the register file without scratch spills: 2 rows in
        // 1024 threads, 4 rows in 512, and (FP32 only) 4 rows in 768
        if (rows == 2 && block == 1024) return CFM_W3(2, 4);
        if constexpr (sizeof(T) == 4) {
            if (rows == 4 && block == 768) return CFM_W3(4, 3);
        }
        if (rows == 4 && block == 512) return CFM_W3(4, 2);
        return nullptr;
    }
    const bool wide = block != 256;
    switch (rows) {
        case 4: return wide ? CFM_J2(4, 1024) : CFM_J2(4, 256);
        case 8: return wide ? CFM_J2(8, 768) : CFM_J2(8, 256);
        case 16: return wide ? CFM_J2(16, 512) : CFM_J2(16, 256);
        default: return nullptr;
    }
#undef CFM_J3
#undef CFM_W3
#undef CFM_J2
}

// Two sweeps per pass: FP64 4 x 4 wavefronts (512 columns x 4 rows), FP32
// 2 x 8 (512 columns x 8 rows); both tiles are 512 wide.
constexpr int kT2Width = 512;
// Two sweeps per pass.  FP64: 4 x 4 wavefronts of 2 columns per lane (1024
// threads) or, with 512 threads, 2 x 4 wavefronts of 4 columns per lane
// (twice the registers per thread, half the per-wave overhead per point).
// FP32: 2 x 8 wavefronts of 4 columns per lane.  All tiles are 512 wide.
template <typename T, int WX, int WY, int VXT, bool TM, bool FULL_OK = true, bool PER = false>
static KernelFn pick_t2_shape(bool aligned, bool full) {
    if constexpr (FULL_OK) {
        if (full) return (KernelFn)jacobi3d_t2_kernel<T, WX, WY, VXT, true, true, PER, TM>;
    }
    return aligned ? (KernelFn)jacobi3d_t2_kernel<T, WX, WY, VXT, true, false, PER, TM>
                   : (KernelFn)jacobi3d_t2_kernel<T, WX, WY, VXT, false, false, PER, TM>;
}
// periodic z: built for the default 512-thread shapes only (null otherwise)
template <typename T, bool TM>
static KernelFn pick_t2_kernel(bool aligned, bool full, int block, bool periodic, bool xt, bool big) {
    // rows wider than one tile (x tiles) and fields of 2 GiB and more (plane windows): the
    // 512-thread buffer-load builds
    if (xt || big) {
        if (!aligned || block != 512) return nullptr;
        if constexpr (sizeof(T) == 8) {
            if (xt)
                return periodic ? (KernelFn)jacobi3d_t2_kernel<T, 2, 4, 4, true, false, true, TM, true, true>
                                : (KernelFn)jacobi3d_t2_kernel<T, 2, 4, 4, true, false, false, TM, true, true>;
            return periodic ? (KernelFn)jacobi3d_t2_kernel<T, 2, 4, 4, true, false, true, TM, false, true>
                            : (KernelFn)jacobi3d_t2_kernel<T, 2, 4, 4, true, false, false, TM, false, true>;
        } else {
            if (xt)
                return periodic ? (KernelFn)jacobi3d_t2_kernel<T, 1, 8, 8, true, false, true, TM, true, true>
                                : (KernelFn)jacobi3d_t2_kernel<T, 1, 8, 8, true, false, false, TM, true, true>;
            return periodic ? (KernelFn)jacobi3d_t2_kernel<T, 1, 8, 8, true, false, true, TM, false, true>
                            : (KernelFn)jacobi3d_t2_kernel<T, 1, 8, 8, true, false, false, TM, false, true>;
        }
    }
    if constexpr (sizeof(T) == 8) {
        if (block == 512)
            return periodic ? pick_t2_shape<T, 2, 4, 4, TM, true, true>(aligned, full)
                            : pick_t2_shape<T, 2, 4, 4, TM>(aligned, full);
        if (periodic) return nullptr;
        // the full-width (nx = 512) build of this shape spills to scratch; the general
        // aligned build handles nx = 512 as well
        return pick_t2_shape<T, 4, 4, 2, TM, false>(aligned, full);
    } else {
        // (the full-width periodic build of this shape spills to scratch: the general aligned one
        // serves nx = 512 too)
        if (block == 512)
            return periodic ? pick_t2_shape<T, 1, 8, 8, TM, false, true>(aligned, full)
                            : pick_t2_shape<T, 1, 8, 8, TM>(aligned, full);
        if (periodic) return nullptr;
        // (the lookahead builds of this shape spill to scratch: it keeps the element-load build)
        return (KernelFn)jacobi3d_t2_kernel<T, 2, 8, 4, false, false, false, TM>;
    }
}

template <typename T, bool TM>
static KernelFn pick_kernel(bool is3d, int rows, bool aligned, int block) {
    return aligned ? pick_kernel_t<T, true, TM>(is3d, rows, block) : pick_kernel_t<T, false, TM>(is3d, rows, block);
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
    JacobiConfig c{};
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;  // buffer holding the current field
    GridSync *sync = nullptr;
    double *norm_part = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long *timers = nullptr;  // phase_timers: [grid][kTimerSlots] ticks
    double clock_khz = 100000.0;           // s_memrealtime rate
    cfm_phase_times phases{};              // breakdown of the last run
};

static cfm_status validate(const cfm_jacobi_params *p) {
    if (!p) return CFM_EINVAL;
    if (p->nx < 3 || p->ny < 3 || p->nz < 1 || p->nz == 2) return CFM_EINVAL;
    if (p->niter < 0) return CFM_EINVAL;
    if (p->prec != CFM_F64 && p->prec != CFM_F32) return CFM_EINVAL;
    if (p->bc != CFM_BC_PERIODIC && p->bc != CFM_BC_DIRICHLET) return CFM_EINVAL;
    if (p->variant < CFM_JACOBI_LAUNCH_PER_ITER || p->variant > CFM_JACOBI_NO_COMPUTE)
        return CFM_EINVAL;
    if (p->time_block < 0 || p->time_block > 4) return CFM_EINVAL;
    // 32-bit plane/row arithmetic inside the kernels
    const double elems = (double)p->nx * p->ny * p->nz;
    if (elems > 9.0e15 || (double)p->nx * p->ny > 2.0e9) return CFM_EINVAL;
    return CFM_OK;
}

cfm_status cfm::jacobi_configure(const cfm_jacobi_params *p, JacobiConfig *c) {
    cfm_status st = validate(p);
    if (st != CFM_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CFM_EDEVICE;
    st = current_device_props(&c->dev);
    if (st != CFM_OK) return st;
    const bool is3d = p->nz >= 3;
    const bool f64 = p->prec == CFM_F64;
    c->elem_size = f64 ? 8 : 4;
    c->elems = (size_t)p->nx * p->ny * (size_t)p->nz;
    const int vx = 16 / (int)c->elem_size;
    const bool aligned = ((size_t)p->nx * c->elem_size) % 16 == 0;
    // the timed build (phase_timers) of a persistent kernel; one launch per sweep is timed by the caller
    const bool tm = p->phase_timers != 0 && p->variant != CFM_JACOBI_LAUNCH_PER_ITER;
    // measured defaults on MI355X: 3D FP64 rows 2 x 1024 threads, 3D FP32 rows 4 x 768, 2D rows 4
    // two sweeps per pass: persistent Dirichlet; 3D with one tile across x
    // (3D periodic: 512-thread workgroups only)
    // the buffer-load (ring) build of the 3D two-sweep kernel reads a unit's planes through a
    // window of chunk_len + 6 planes with 32-bit offsets: 16-byte aligned rows, planes small enough
    // for windows of at least 9 (chunk lengths are capped below to keep the window under 2 GiB)
    const size_t plane_bytes = (size_t)p->nx * p->ny * c->elem_size;
    const bool ring_ok = aligned && 9 * plane_bytes < 0x80000000ull;
    // rows wider than one tile (x tiles): the ring build in 512-thread workgroups
    const bool xt = is3d && p->nx > kT2Width;
    const bool xt_ok = ring_ok && (p->block_threads <= 0 || p->block_threads == 512);
    const bool t2_ok3 = is3d && p->variant == CFM_JACOBI_PERSISTENT && (!xt || xt_ok) && p->tile_rows <= 0 &&
                        (p->bc == CFM_BC_DIRICHLET
                             ? (p->block_threads <= 0 || p->block_threads == 1024 || p->block_threads == 512)
                             : (p->block_threads <= 0 || p->block_threads == 512));
    // 2D: two sweeps per pass with Dirichlet or the reference's periodic y rows, three and four
    // with Dirichlet rows
    const bool t2_ok2 = !is3d && p->variant == CFM_JACOBI_PERSISTENT && p->tile_rows <= 0 &&
                        (p->block_threads <= 0 || p->block_threads == 1024 || p->block_threads == 512);
    const bool t34_ok2 = t2_ok2 && p->bc == CFM_BC_DIRICHLET;
    const bool t2_ok = t2_ok3 || t2_ok2;
    if (p->time_block == 2 && !t2_ok)
        return CFM_FAIL(CFM_ENOTSUP, "jacobi: two sweeps per pass needs the persistent variant, the default tile "
                                     "shape and, in 3D, 512-thread workgroups for periodic z and, with nx > %d, "
                                     "16-byte aligned rows and planes below 2 GiB / 9", kT2Width);
    if (p->time_block >= 3 && !t34_ok2)
        return CFM_FAIL(CFM_ENOTSUP, "jacobi: three or four sweeps per pass need the persistent 2D Dirichlet "
                                     "variant and the default tile shape");
    if (p->time_block == 4 && p->block_threads > 0 && p->block_threads != 512)
        return CFM_FAIL(CFM_ENOTSUP, "jacobi: four sweeps per pass run in 512-thread workgroups");
    // x tiles compute kT2Width columns per tile and own kT2Width - vx of them: auto takes them only
    // when the tiles' owned share of their columns is at least 0.70 (two sweeps per pass over one,
    // FP64: 1022 and 2042 columns (share 1.0) 1.47 x and 1.51 x, 2048 (0.80) 1.09 x, 768 (0.75)
    // 1.32 x, 1536 (0.75) 1.04 x, 1024 (0.67) 0.98 x; FP32 2048 1.05 x, 1536 1.02 x, 1024 0.91 x)
    const int xs = kT2Width - vx;
    const double xshare = xt ? (double)(p->nx - 2) / ((double)((p->nx - 2 + xs - 1) / xs) * xs) : 1.0;
    c->t2 = t2_ok && p->time_block != 1 && !(xt && p->time_block == 0 && xshare < 0.70);
    // 2D: four sweeps per pass by default with the default workgroup (16384^2 FP64 976 GLUP/s vs 790
    // for three and 617 for two; FP32 1612 vs 1523 and 1206); three when 1024 threads are asked for
    // (periodic y rows: two)
    const int auto2d = p->bc != CFM_BC_DIRICHLET ? 2 : (p->block_threads <= 0 || p->block_threads == 512) ? 4 : 3;
    c->sweeps_per_pass =
        !c->t2 ? 1 : p->time_block >= 3 ? p->time_block : (p->time_block == 0 && !is3d) ? auto2d : 2;
    if (c->t2 && !is3d) {
        c->rows = 1;
        const bool per2 = p->bc == CFM_BC_PERIODIC;
        // 16384^2: FP64 617 GLUP/s at 1024 threads (2 columns per lane) vs 534 at 512;
        // FP32 1205 at 1024 threads vs 942 at 512
        const bool wide_ok = !(c->sweeps_per_pass == 3 && f64 && !aligned);  // (that build spills at 1024)
        c->block = p->block_threads > 0 ? p->block_threads : (c->sweeps_per_pass == 4 || !wide_ok ? 512 : 1024);
        c->kernel = c->sweeps_per_pass == 4 ? (tm ? pick_t4d_kernel<true>(f64, aligned)
                                                  : pick_t4d_kernel<false>(f64, aligned))
                    : c->sweeps_per_pass == 3 ? (tm ? pick_t3d_kernel<true>(f64, aligned, c->block)
                                                    : pick_t3d_kernel<false>(f64, aligned, c->block))
                                              : (tm ? pick_t2d_kernel<true>(f64, aligned, c->block, per2)
                                                    : pick_t2d_kernel<false>(f64, aligned, c->block, per2));
    } else if (c->t2) {
        c->rows = 1;
        // 512-thread workgroups measured best at 512^3: FP64 4 columns per lane 478 vs 428
        // GLUP/s (2 columns, 1024 threads); FP32 one wavefront across x (8 columns per lane)
        // 882 vs 844 (4 columns, 1024 threads)
        // 16-byte aligned rows use buffer loads with 32-bit offsets into each unit's window of
        // planes (the two-step load lookahead, 512^3 FP64 477 -> 545 GLUP/s)
        c->block = p->block_threads > 0 ? p->block_threads : 512;
        // fields of 2 GiB and more: the 512-thread ring builds with plane windows, else the
        // element-load build (1024 threads)
        const bool big = c->elems * c->elem_size >= 0x80000000ull;
        const bool al = ring_ok && (!big || c->block == 512);
        const bool full = p->nx == kT2Width && al;
        const bool per = p->bc == CFM_BC_PERIODIC;
        c->kernel = f64 ? (tm ? pick_t2_kernel<double, true>(al, full && !big, c->block, per, xt, big && al)
                              : pick_t2_kernel<double, false>(al, full && !big, c->block, per, xt, big && al))
                        : (tm ? pick_t2_kernel<float, true>(al, full && !big, c->block, per, xt, big && al)
                              : pick_t2_kernel<float, false>(al, full && !big, c->block, per, xt, big && al));
    } else {
        c->rows = p->tile_rows > 0 ? p->tile_rows : (is3d ? (f64 ? 2 : 4) : 4);
        c->block = p->block_threads > 0 ? p->block_threads : default_block(is3d, c->rows, f64);
        if (!is3d && c->block != 256 && c->block != default_block(false, c->rows, f64))
            return CFM_FAIL(CFM_EINVAL, "jacobi: 2D %d-row tiles run in 256 or %d-thread workgroups", c->rows,
                            default_block(false, c->rows, f64));
        c->kernel = f64 ? (tm ? pick_kernel<double, true>(is3d, c->rows, aligned, c->block)
                              : pick_kernel<double, false>(is3d, c->rows, aligned, c->block))
                        : (tm ? pick_kernel<float, true>(is3d, c->rows, aligned, c->block)
                              : pick_kernel<float, false>(is3d, c->rows, aligned, c->block));
    }
    if (!c->kernel)
        return CFM_FAIL(CFM_EINVAL, "jacobi: no kernel for %d-row tiles in %d-thread workgroups (%s, %s rows)",
                        c->rows, c->block, f64 ? "FP64" : "FP32", aligned ? "16-byte aligned" : "unaligned");
    // grid: all co-resident workgroups
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)c->kernel, c->block, 0) !=
            hipSuccess ||
        per_cu < 1)
        return CFM_EDEVICE;
    // the barrier keeps one arrival word per workgroup (kMaxGrid of them)
    const int max_grid = std::min(per_cu * c->dev.cus, kMaxGrid);
    c->grid = p->grid_blocks > 0 ? std::min(p->grid_blocks, max_grid) : max_grid;
    // work units: one per wavefront (per-wave kernels) or per workgroup (3D tile kernel)
    const bool wg_tile = is3d && c->block != 256;
    const int nworkers = wg_tile ? c->grid : c->grid * (c->block / 64);
    int tile_w = wg_tile ? 64 * vx * 4 : 64 * vx;
    if (c->t2 && !is3d) tile_w = 64 * (f64 ? t2d_cols<double>(c->block) : t2d_cols<float>(c->block));
    int tile_h = wg_tile ? c->rows * (c->block / 256) : c->rows;
    if (c->t2 && is3d) {
        tile_w = kT2Width;
        tile_h = f64 ? 4 : 8;
    }
    c->ntx = (p->nx + tile_w - 1) / tile_w;
    // x tiles of the two-sweep 3D kernel: stride kT2Width - vx, each owning the columns after its
    // first (the edge columns' level 2 needs level 1 one column further out)
    if (c->t2 && is3d) c->ntx = xt ? (p->nx - 2 + (kT2Width - vx) - 1) / (kT2Width - vx) : 1;
    const int interior = is3d ? p->nz - 2 : p->ny - 2;
    c->nty = is3d ? (p->ny + tile_h - 1) / tile_h : 1;
    const int tiles = c->ntx * c->nty;
    int chunks = p->chunks;
    // the ring build's window of chunk_len + 6 planes stays below 2 GiB: the fewest chunks that allows
    int min_chunks = 1;
    if (c->t2 && is3d && ring_ok && (xt || c->elems * c->elem_size >= 0x80000000ull)) {
        const long long max_len = std::max(3LL, (long long)(0x7fffffffull / plane_bytes) - 6);
        min_chunks = (int)((interior + max_len - 1) / max_len);
    }
    if (chunks > 0) chunks = std::max(chunks, min_chunks);
    if (chunks <= 0 && c->t2 && is3d) {
        // two sweeps per pass in 3D: the chunk count whose units spread best over the grid
        // (units / (grid x rounds)), each chunk paying two warm-up planes; 512^3 FP64: 128 tiles,
        // 2 chunks = one unit per workgroup; 1024 x 512 x 256: 384 tiles, 2 chunks = 3 rounds
        double best = -1.0;
        for (int ch = min_chunks; ch <= std::max(min_chunks, std::min(16, interior / 16)); ++ch) {
            const long long units = (long long)tiles * ch;
            const long long rounds = (units + c->grid - 1) / c->grid;
            const double len = (double)interior / ch;
            const double eff = (double)units / ((double)rounds * c->grid) * len / (len + 2.0);
            if (eff > best + 1e-9) {
                best = eff;
                chunks = ch;
            }
        }
    } else if (chunks <= 0) {
        // 3D tiles of 2 rows: 4 z-chunks per workgroup balance best; else one
        const int per_worker = (wg_tile && !c->t2 && c->rows <= 2) ? 4 : 1;
        chunks = std::max(1, per_worker * nworkers / tiles);
        const int min_len = is3d ? 16 : 64;
        chunks = std::min(chunks, std::max(1, interior / min_len));
    }
    chunks = std::min(chunks, interior);
    c->chunk_len = (interior + chunks - 1) / chunks;
    c->chunks = (interior + c->chunk_len - 1) / c->chunk_len;
    c->norm_slots = p->norm_every > 0 ? (p->niter + p->norm_every - 1) / p->norm_every : 0;
    c->spin_ticks = spin_ticks(p->spin_limit_us);
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->dev.device) ==
            hipSuccess &&
        rate_khz > 0) {
        const unsigned long long us = p->spin_limit_us > 0 ? p->spin_limit_us : 10000000ull;
        c->spin_ticks = us * (unsigned long long)rate_khz / 1000ull;
    }
    return CFM_OK;
}

cfm_status cfm::jacobi_enqueue(const cfm_jacobi_params &p, const JacobiConfig &c, void *b0, void *b1,
                               GridSync *sync, double *norm_part, unsigned long long *timers, hipStream_t stream,
                               int *coop) {
    JacobiArgs A{};
    A.buf0 = b0;
    A.buf1 = b1;
    A.nx = p.nx;
    A.ny = p.ny;
    A.nz = p.nz;
    A.bc = p.bc;
    A.norm_every = p.norm_every;
    A.chunks = c.chunks;
    A.chunk_len = c.chunk_len;
    A.ntx = c.ntx;
    A.nty = c.nty;
    A.norm_part = norm_part;
    A.sync = sync;
    A.timers = p.variant == CFM_JACOBI_LAUNCH_PER_ITER ? nullptr : timers;
    A.no_compute = p.variant == CFM_JACOBI_NO_COMPUTE;
    *coop = -1;
    hipLaunchKernelGGL(reset_sync_kernel, dim3(8), dim3(256), 0, stream, sync, c.spin_ticks, c.grid);
    CFM_HIP(hipGetLastError());
    if (p.niter <= 0) return CFM_OK;
    if (p.variant == CFM_JACOBI_LAUNCH_PER_ITER) {
        for (int k = 0; k < p.niter; ++k) {
            A.niter = 1;
            A.it0 = k;
            hipLaunchKernelGGL(c.kernel, dim3(c.grid), dim3(c.block), 0, stream, A);
            CFM_HIP(hipGetLastError());
        }
    } else {
        A.niter = p.niter;
        A.it0 = 0;
        void *args[] = {&A};
        CFM_HIP(launch_persistent((const void *)c.kernel, c.grid, c.block, args, stream, coop));
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_create(const cfm_jacobi_params *p, cfm_jacobi_plan **out) {
    if (!out) return CFM_EINVAL;
    *out = nullptr;
    JacobiConfig cfg;
    cfm_status st = jacobi_configure(p, &cfg);
    if (st != CFM_OK) return st;
    auto *pl = new (std::nothrow) cfm_jacobi_plan();
    if (!pl) return CFM_ENOMEM;
    pl->p = *p;
    pl->c = cfg;
    auto fail = [&](cfm_status s) {
        cfm_jacobi_plan_destroy(pl);
        return s;
    };
    if (hipStreamCreate(&pl->stream) != hipSuccess) return fail(CFM_EDEVICE);
    if (hipEventCreate(&pl->ev0) != hipSuccess || hipEventCreate(&pl->ev1) != hipSuccess)
        return fail(CFM_EDEVICE);
    for (int i = 0; i < 2; ++i)
        if (hipMalloc(&pl->buf[i], pl->c.elems * pl->c.elem_size) != hipSuccess)
            return fail(CFM_ENOMEM);
    if (hipMalloc((void **)&pl->sync, sizeof(GridSync)) != hipSuccess) return fail(CFM_ENOMEM);
    if (pl->c.norm_slots > 0 &&
        hipMalloc((void **)&pl->norm_part, sizeof(double) * pl->c.norm_slots * pl->c.grid) != hipSuccess)
        return fail(CFM_ENOMEM);
    if (p->phase_timers && p->variant != CFM_JACOBI_LAUNCH_PER_ITER) {
        if (hipMalloc((void **)&pl->timers, sizeof(unsigned long long) * kTimerSlots * pl->c.grid) != hipSuccess)
            return fail(CFM_ENOMEM);
        int rate_khz = 0;
        if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, pl->c.dev.device) == hipSuccess &&
            rate_khz > 0)
            pl->clock_khz = rate_khz;
    }
    *out = pl;
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_load(cfm_jacobi_plan *pl, const void *init_host) {
    if (!pl || !init_host) return CFM_EINVAL;
    const size_t bytes = pl->c.elems * pl->c.elem_size;
    CFM_HIP(hipMemcpyAsync(pl->buf[0], init_host, bytes, hipMemcpyHostToDevice, pl->stream));
    CFM_HIP(hipMemcpyAsync(pl->buf[1], pl->buf[0], bytes, hipMemcpyDeviceToDevice, pl->stream));
    CFM_HIP(hipStreamSynchronize(pl->stream));
    pl->cur = 0;
    return CFM_OK;
}

int cfm::jacobi_passes(const cfm_jacobi_params &p, const JacobiConfig &c) {
    if (p.variant == CFM_JACOBI_NO_COMPUTE) return 0;
    return (p.niter + c.sweeps_per_pass - 1) / c.sweeps_per_pass;
}

void cfm::jacobi_fill_timing(cfm_timing *t, const cfm_jacobi_params &p, const JacobiConfig &c, float ms,
                             double wall_ms, int coop) {
    if (!t) return;
    t->kernel_ms = ms;
    t->wall_ms = wall_ms;
    t->grid_blocks = c.grid;
    t->block_threads = c.block;
    t->iterations = p.niter;
    t->outer_iterations = jacobi_passes(p, c);
    t->cooperative = coop;
}

static cfm_status jacobi_run(cfm_jacobi_plan *pl, double *norms, int norm_cap, int *norm_count, cfm_timing *t);

// A roctx range around every run (a no-op unless a profiler listens).
extern "C" cfm_status cfm_jacobi_plan_run(cfm_jacobi_plan *pl, double *norms, int norm_cap,
                                          int *norm_count, cfm_timing *t) {
    if (!pl) return CFM_EINVAL;
    roctxRangePush(pl->p.nz >= 3 ? "cfm_jacobi_plan_run 3D" : "cfm_jacobi_plan_run 2D");
    const cfm_status st = jacobi_run(pl, norms, norm_cap, norm_count, t);
    roctxRangePop();
    return st;
}

static cfm_status jacobi_run(cfm_jacobi_plan *pl, double *norms, int norm_cap, int *norm_count, cfm_timing *t) {
    const auto w0 = std::chrono::steady_clock::now();
    const cfm_jacobi_params &p = pl->p;
    const JacobiConfig &c = pl->c;
    int coop = -1;
    CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
    cfm_status st = jacobi_enqueue(p, c, pl->buf[pl->cur], pl->buf[pl->cur ^ 1], pl->sync, pl->norm_part,
                                   pl->timers, pl->stream, &coop);
    if (st != CFM_OK) return st;
    CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
    CFM_HIP(hipEventSynchronize(pl->ev1));
    float ms = 0.f;
    CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
    GridSync hs{};
    CFM_HIP(hipMemcpy(&hs, pl->sync, sizeof(hs), hipMemcpyDeviceToHost));
    if (hs.abort) return barrier_timeout("jacobi", coop);
    pl->phases = cfm_phase_times{};
    if (pl->timers && p.niter > 0) {
        std::vector<unsigned long long> ticks((size_t)kTimerSlots * c.grid);
        CFM_HIP(hipMemcpy(ticks.data(), pl->timers, sizeof(unsigned long long) * ticks.size(), hipMemcpyDeviceToHost));
        summarize_phases(ticks.data(), c.grid, pl->clock_khz, &pl->phases);
    }
    int produced = 0;
    if (c.norm_slots > 0 && p.variant != CFM_JACOBI_NO_COMPUTE) {
        std::vector<double> part((size_t)c.norm_slots * c.grid);
        CFM_HIP(hipMemcpy(part.data(), pl->norm_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
        for (int s = 0; s < c.norm_slots; ++s) {
            double acc = 0.0;
            for (int b = 0; b < c.grid; ++b) acc += part[(size_t)s * c.grid + b];
            if (norms && s < norm_cap) norms[s] = std::sqrt(acc);
        }
        produced = c.norm_slots;
    }
    if (norm_count) *norm_count = produced;
    pl->cur ^= jacobi_passes(p, c) & 1;
    jacobi_fill_timing(t, p, c, ms,
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count(), coop);
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_phase_times(cfm_jacobi_plan *pl, cfm_phase_times *out) {
    if (!pl || !out) return CFM_EINVAL;
    if (!pl->timers)
        return CFM_FAIL(CFM_EINVAL, "jacobi: the plan was created without phase_timers (or runs one launch per sweep)");
    *out = pl->phases;
    return CFM_OK;
}

extern "C" cfm_status cfm_jacobi_plan_download(cfm_jacobi_plan *pl, void *out_host) {
    if (!pl || !out_host) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(out_host, pl->buf[pl->cur], pl->c.elems * pl->c.elem_size, hipMemcpyDeviceToHost));
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
    if (pl->timers) (void)hipFree(pl->timers);
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
