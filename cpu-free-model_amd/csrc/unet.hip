// Learned warm start: a 3-level 3D U-Net (16/32/64 channels, 3x3x3 convs)
// with hand-written forward and backward kernels on BF16 MFMA (FP32
// accumulate), trained on the GPU with Adam.  No counterpart in the
// reference: it produces initial guesses x0 for the CG solvers.
//
// Tensors are channels-last (NDHWC) BF16; spatial index p = ((b*n+z)*n+y)*n+x.
// A 3x3x3 convolution is an implicit GEMM: Y[p][co] = sum_k W[co][k] Xs[p][k]
// with k = tap*Cin + ci and Xs[p][k] = X[p + offset(tap)][ci] (zero outside).
// One v_mfma_f32_16x16x32_bf16 handles 16 output channels x 16 points x 32
// k; lane l holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and
// D[4(l/16)+i][l%16] (layout checked on the device).  Channel counts are
// multiples of 8, so a lane's 8 k-values are 8 consecutive channels of one
// tap: one 16-byte read from the workgroup's input block, which is staged in
// LDS with its one-point halo.
//
// Backward data is the same kernel with the weights flipped in space and
// transposed in channels; backward weights is a split-K MFMA GEMM over point
// slabs (one workgroup per tap and slab, point blocks staged transposed in
// LDS) followed by a fixed-order reduction of the slab partials.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cfm_common.hpp"
#include "unet.hpp"

namespace cfm {
namespace unet {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2bf(float v) { return (bf16)v; }  // round to nearest even

__device__ __forceinline__ int tap_off(int tap, int &dz, int &dy, int &dx) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
    return 0;
}

// ------------------------------------------------------------- conv fwd ---
// Y = conv3x3x3(X) + bias, optional ReLU.  W is [Cout][Kp] BF16 (k = tap*Cin+ci,
// zero padded to Kp, a multiple of 32).
// LDS-staged forward conv for the layer shapes of the network.  A workgroup
// owns 16 (x) x TY (y) x TZ (z) output points of one batch element; the input
// block with its one-point halo and all CIN channels is staged in LDS once
// (zeros outside the grid), so a lane's B fragment (8 channels of one tap at
// one point) is a single ds_read_b128 with no bounds checks.  Each of the
// four wavefronts computes NT = TY*TZ/4 rows of 16 points (rows w*NT .. of the
// tile, y fastest) for all MT channel tiles: every weight fragment (read from
// the L1/L2-resident weights) feeds NT MFMAs and every input fragment MT.
// The four accumulators of a lane (output channels 4g..4g+3 of one point) are
// biased, rectified, rounded to BF16 and written as one 8-byte store.
template <int MT, int CIN, int TY, int TZ>
__global__ void __launch_bounds__(256) conv3_fwd_lds_kernel(const bf16 *__restrict__ X, const bf16 *__restrict__ W,
                                                            int Kp, const float *__restrict__ bias,
                                                            bf16 *__restrict__ Y, int n, int nbx, int nby, int nbz,
                                                            int relu, const bf16 *__restrict__ mask) {
    constexpr int HX = 18, HY = TY + 2, HZ = TZ + 2, NV = CIN / 8;
    constexpr int NT = TY * TZ / 4;
    static_assert(CIN % 8 == 0 && (TY * TZ) % 4 == 0, "shape");
    __shared__ bf16x8 blk[HZ * HY * HX * NV];
    int bid = blockIdx.x;
    const int bx = bid % nbx;
    bid /= nbx;
    const int by = bid % nby;
    bid /= nby;
    const int bz = bid % nbz;
    const int b = bid / nbz;
    const int x0 = bx * 16, y0 = by * TY, z0 = bz * TZ;
    const bf16x8 zero8 = {(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f,
                          (bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
    for (int v = threadIdx.x; v < HZ * HY * HX * NV; v += 256) {
        const int cv = v % NV;
        int r = v / NV;
        const int cx = r % HX;
        r /= HX;
        const int cy = r % HY, cz = r / HY;
        const int gx = x0 - 1 + cx, gy = y0 - 1 + cy, gz = z0 - 1 + cz;
        bf16x8 val = zero8;
        if (gx >= 0 && gx < n && gy >= 0 && gy < n && gz >= 0 && gz < n)
            val = *reinterpret_cast<const bf16x8 *>(X + ((((size_t)b * n + gz) * n + gy) * n + gx) * CIN + cv * 8);
        blk[v] = val;
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nl = lane & 15, g = lane >> 4;
    int ty[NT], tz[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        ty[t] = (w * NT + t) % TY;
        tz[t] = (w * NT + t) / TY;
    }
    constexpr int K = 27 * CIN, KPC = (K + 31) / 32 * 32;  // Kp == KPC (checked by the host)
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // a compile-time trip count, unrolled by eight: the weight loads of eight k steps go out
    // together instead of one L2 round trip per k step
#pragma unroll 8
    for (int kb = 0; kb < KPC; kb += 32) {
        const int k = kb + 8 * g;
        bf16x8 av[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const bf16x8 *>(W + (size_t)(m * 16 + nl) * Kp + k);
        const bool kin = k < K;
        const int tap = kin ? k / CIN : 26;  // the zero-padded tail of K reads a valid slot, then discards it
        const int ci8 = kin ? (k % CIN) / 8 : 0;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // offsets + 1
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            bf16x8 bv = blk[(((tz[t] + dz) * HY + (ty[t] + dy)) * HX + (nl + dx)) * NV + ci8];
            if (!kin) bv = zero8;
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv, acc[m][t], 0, 0, 0);
        }
    }
    // D[4g+i][nl]: output channel m*16 + 4g + i of point (x0 + nl, y0 + ty[t], z0 + tz[t])
    constexpr int Cout = MT * 16;
    const int x = x0 + nl;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int y = y0 + ty[t], z = z0 + tz[t];
        if (x >= n || y >= n || z >= n) continue;
        const size_t p = (((size_t)b * n + z) * n + y) * n + x;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            bf16x4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[m][t][i] + (bias ? bias[m * 16 + 4 * g + i] : 0.f);
                if (relu && v < 0.f) v = 0.f;
                o[i] = f2bf(v);
            }
            if (mask) {  // ReLU backward: zero where the forward activation was not positive
                const bf16x4 mk = *reinterpret_cast<const bf16x4 *>(mask + p * Cout + m * 16 + 4 * g);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (!(bf2f(mk[i]) > 0.f)) o[i] = (bf16)0.0f;
            }
            *reinterpret_cast<bf16x4 *>(Y + p * Cout + m * 16 + 4 * g) = o;
        }
    }
}

// Forward conv with the three x taps in three accumulators, for rows of at
// most 64 points.  A workgroup owns whole x rows (NXT 16-point tiles, zeros
// beyond n) of TY x TZ (y, z) rows of one batch element and stages their
// (TY+2) x (TZ+2) input rows, all CIN channels, in LDS; no x halo is needed.
// K runs over the nine (dz, dy) taps and the channels, k = (3 dz + dy) CIN + ci
// (a lane's 8 values: 8 channels of one tap).  Per 32-wide k block a lane reads
// ONE input fragment, X at its own point x, rows y + dy, z + dz, and feeds it to
// the three dx taps with their own weight fragments W[co][(3 (3 dz + dy) + dx) CIN
// + ci]: acc[dx] at x collects the terms of X[x] that belong to output
// Y[x + 1 - dx].  After the k loop Y[x] = acc[1][x] + acc[0][x - 1] + acc[2][x + 1]
// (zero past the row ends: the padding there is zero): a DPP row shift per value,
// the lane across a 16-point tile edge taken from the neighbouring tile by a row
// rotation.  Per input fragment read from LDS the wavefront issues 3 MT MFMAs,
// against MT when the x taps are separate k values.  RPW rows per wavefront; the
// k loop unrolled by KU.
template <int MT, int CIN, int TY, int TZ, int NXT, int NW, int KU>
__global__ void __launch_bounds__(64 * NW) conv3_fwd_xacc_kernel(const bf16 *__restrict__ X,
                                                                 const bf16 *__restrict__ W, int Kp,
                                                                 const float *__restrict__ bias,
                                                                 bf16 *__restrict__ Y, int n, int nby, int nbz,
                                                                 int relu, const bf16 *__restrict__ mask) {
    constexpr int XW = NXT * 16, HY = TY + 2, HZ = TZ + 2, NV = CIN / 8;
    constexpr int RPW = TY * TZ / NW;
    constexpr int K9 = 9 * CIN, KB = (K9 + 31) / 32;
    static_assert(CIN % 8 == 0 && (TY * TZ) % NW == 0, "shape");
    __shared__ bf16x8 blk[HZ * HY * XW * NV];
    int bid = blockIdx.x;
    const int by = bid % nby;
    bid /= nby;
    const int bz = bid % nbz;
    const int b = bid / nbz;
    const int y0 = by * TY, z0 = bz * TZ;
    const bf16x8 zero8 = {(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f,
                          (bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
    for (int v = threadIdx.x; v < HZ * HY * XW * NV; v += 64 * NW) {
        const int cv = v % NV;
        int r = v / NV;
        const int cx = r % XW;
        r /= XW;
        const int cy = r % HY, cz = r / HY;
        const int gy = y0 - 1 + cy, gz = z0 - 1 + cz;
        bf16x8 val = zero8;
        if (cx < n && gy >= 0 && gy < n && gz >= 0 && gz < n)
            val = *reinterpret_cast<const bf16x8 *>(X + ((((size_t)b * n + gz) * n + gy) * n + cx) * CIN + cv * 8);
        blk[v] = val;
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nl = lane & 15, g = lane >> 4;
    int ry[RPW], rz[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        ry[r] = (w * RPW + r) % TY;
        rz[r] = (w * RPW + r) / TY;
    }
    f32x4 acc[3][MT][RPW][NXT];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int t = 0; t < NXT; ++t) acc[d][m][r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16 *wrow = W + (size_t)nl * Kp;
#pragma unroll KU
    for (int kb = 0; kb < KB; ++kb) {
        const int k = kb * 32 + 8 * g;
        // (K9 not a multiple of 32: the tail lanes read tap 8's fragment, then zero it)
        const bool kin = K9 % 32 == 0 || k < K9;
        const int dzdy = kin ? k / CIN : 8;
        const int ci = kin ? k % CIN : 0;
        const int dz = dzdy / 3, dy = dzdy % 3;
        bf16x8 av[3][MT];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < MT; ++m)
                av[d][m] = *reinterpret_cast<const bf16x8 *>(wrow + (size_t)m * 16 * Kp + (dzdy * 3 + d) * CIN + ci);
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int t = 0; t < NXT; ++t) {
                bf16x8 bv = blk[(((rz[r] + dz) * HY + (ry[r] + dy)) * XW + t * 16 + nl) * NV + ci / 8];
                if (!kin) bv = zero8;
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int m = 0; m < MT; ++m)
                        acc[d][m][r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[d][m], bv, acc[d][m][r][t], 0, 0, 0);
            }
    }
    // D[4g+i][nl] of tile t: output channel m*16 + 4g + i at x = 16 t + nl
    constexpr int Cout = MT * 16;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int y = y0 + ry[r], z = z0 + rz[r];
#pragma unroll
        for (int t = 0; t < NXT; ++t) {
            const int x = t * 16 + nl;
            const bool out = x < n && y < n && z < n;
            const size_t p = (((size_t)b * n + z) * n + y) * n + x;
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // (each element copied to a float first: a bit cast of a vector element
                    // subscript reads element 0 in this compiler)
                    const float c0 = acc[0][m][r][t][i], c2 = acc[2][m][r][t][i];
                    const float p0 = t > 0 ? acc[0][m][r][t > 0 ? t - 1 : 0][i] : 0.f;
                    const float q2 = t + 1 < NXT ? acc[2][m][r][t + 1 < NXT ? t + 1 : t][i] : 0.f;
                    // acc[0] one point to the left: lane nl-1 of this tile (row_shr:1), lane 15 of
                    // tile t-1 for nl = 0 (row_ror:1 brings lane 15 of each 16-lane row to lane 0)
                    const int l0 = t > 0 ? __builtin_amdgcn_mov_dpp(__float_as_int(p0), 0x121 /* row_ror:1 */, 0xf, 0xf,
                                                                    false)
                                         : 0;
                    const int s0 = __builtin_amdgcn_update_dpp(l0, __float_as_int(c0), 0x111 /* row_shr:1 */, 0xf,
                                                               0xf, false);
                    // acc[2] one point to the right: lane nl+1 (row_shl:1), lane 0 of tile t+1 for
                    // nl = 15 (row_ror:15 brings lane 0 to lane 15)
                    const int l2 = t + 1 < NXT ? __builtin_amdgcn_mov_dpp(__float_as_int(q2), 0x12f /* row_ror:15 */,
                                                                          0xf, 0xf, false)
                                               : 0;
                    const int s2 = __builtin_amdgcn_update_dpp(l2, __float_as_int(c2), 0x101 /* row_shl:1 */, 0xf,
                                                               0xf, false);
                    float v = acc[1][m][r][t][i] + __int_as_float(s0) + __int_as_float(s2);
                    v += bias ? bias[m * 16 + 4 * g + i] : 0.f;
                    if (relu && v < 0.f) v = 0.f;
                    o[i] = f2bf(v);
                }
                if (out) {
                    if (mask) {  // ReLU backward: zero where the forward activation was not positive
                        const bf16x4 mk = *reinterpret_cast<const bf16x4 *>(mask + p * Cout + m * 16 + 4 * g);
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (!(bf2f(mk[i]) > 0.f)) o[i] = (bf16)0.0f;
                    }
                    *reinterpret_cast<bf16x4 *>(Y + p * Cout + m * 16 + 4 * g) = o;
                }
            }
        }
    }
}

// ------------------------------------------------------ conv weight grad ---
// partial[slab][co][k] = sum over points of slab: dZ[p][co] * Xs[p][k]
// (M = 16 output channels, N = 16 k, K = 32 points per MFMA).  Each
// wavefront owns one (co tile, k tile) pair over one slab of `slab` points.
__global__ void __launch_bounds__(256) conv3_wgrad_kernel(const bf16 *__restrict__ X, int Cin,
                                                          const bf16 *__restrict__ dZ, int Cout, int Kp,
                                                          float *__restrict__ partial, int n, int npts, int slab,
                                                          int ntile_k) {
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ntile = (Cout / 16) * ntile_k;
    const int s = blockIdx.y;
    if (wid >= ntile) return;
    const int mt = wid / ntile_k, kt = wid - mt * ntile_k;
    const int nl = lane & 15, g = lane >> 4;
    const int co = mt * 16 + nl;         // A row (lane's m)
    const int kcol = kt * 16 + nl;       // B column (lane's n)
    const int K = 27 * Cin;
    int tap = 0, ci = 0, dz = 0, dy = 0, dx = 0;
    const bool kvalid = kcol < K;
    if (kvalid) {
        tap = kcol / Cin;
        ci = kcol - tap * Cin;
        tap_off(tap, dz, dy, dx);
    }
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    const int pbeg = s * slab, pend = min(npts, pbeg + slab);
    for (int pb = pbeg; pb < pend; pb += 32) {
        bf16x8 av, bv;
        // coordinates of the lane's first point, then stepped along x with carries
        const int p8 = pb + 8 * g;
        int x = p8 % n, y = (p8 / n) % n, z = (p8 / (n * n)) % n, b = p8 / (n * n * n);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = p8 + j;
            bf16 a = (bf16)0.0f, bb = (bf16)0.0f;
            if (p < pend) {
                a = dZ[(size_t)p * Cout + co];
                if (kvalid) {
                    const int zz = z + dz, yy = y + dy, xx = x + dx;
                    if (zz >= 0 && zz < n && yy >= 0 && yy < n && xx >= 0 && xx < n)
                        bb = X[((((size_t)b * n + zz) * n + yy) * n + xx) * Cin + ci];
                }
            }
            av[j] = a;
            bv[j] = bb;
            if (++x == n) {
                x = 0;
                if (++y == n) {
                    y = 0;
                    if (++z == n) {
                        z = 0;
                        ++b;
                    }
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    // D[4g+i][nl] = dW[mt*16 + 4g + i][kt*16 + nl]
    float *out = partial + (size_t)s * Cout * Kp;
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(mt * 16 + 4 * g + i) * Kp + kt * 16 + nl] = acc[i];
}

// LDS-staged weight gradient for one tap: workgroup (tap, slab) computes
// partial[slab][co][tap*CIN + ci] for all co < COUT, ci < CIN.  Per block of
// 64 consecutive points it stages dZ (64 x COUT) and the tap-shifted input
// (64 x CIN, zero where the neighbour leaves the grid) transposed in LDS,
// [channel][point], so each lane's 8 consecutive points of one channel -- an
// MFMA fragment along the point (K) dimension -- are one ds_read_b128.  The
// (co tile, ci tile) pairs are spread over the four wavefronts.  The tap-26
// workgroups also zero the padding columns [27*CIN, Kp).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3_wgrad_lds_kernel(const bf16 *__restrict__ X, const bf16 *__restrict__ dZ,
                                                              int Kp, float *__restrict__ partial, int n, int npts,
                                                              int slab) {
    constexpr int PB = 64;  // points per staged block
    constexpr int MT = COUT / 16, KT = CIN / 16, T = MT * KT, TW = (T + 3) / 4;
    static_assert(CIN % 16 == 0 && COUT % 16 == 0, "shape");
    __shared__ __attribute__((aligned(16))) bf16 dzT[COUT * PB];
    __shared__ __attribute__((aligned(16))) bf16 xT[CIN * PB];
    const int tap = blockIdx.x, s = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nl = lane & 15, g = lane >> 4;
    f32x4 acc[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f,
                          (bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
    const int pbeg = s * slab, pend = min(npts, pbeg + slab);
    for (int pb = pbeg; pb < pend; pb += PB) {
        // stage dZ[pb..pb+63][*] and the tap-shifted X[pb..pb+63][*], transposed
        for (int v = threadIdx.x; v < PB * (COUT / 8); v += 256) {
            const int j = v / (COUT / 8), c8 = v % (COUT / 8);
            const int p = pb + j;
            const bf16x8 val = p < pend ? *reinterpret_cast<const bf16x8 *>(dZ + (size_t)p * COUT + 8 * c8) : zero8;
#pragma unroll
            for (int i = 0; i < 8; ++i) dzT[(8 * c8 + i) * PB + j] = val[i];
        }
        for (int v = threadIdx.x; v < PB * (CIN / 8); v += 256) {
            const int j = v / (CIN / 8), c8 = v % (CIN / 8);
            const int p = pb + j;
            bf16x8 val = zero8;
            if (p < pend) {
                const int x = p % n, y = (p / n) % n, z = (p / (n * n)) % n, b = p / (n * n * n);
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                if (xx >= 0 && xx < n && yy >= 0 && yy < n && zz >= 0 && zz < n)
                    val = *reinterpret_cast<const bf16x8 *>(
                        X + ((((size_t)b * n + zz) * n + yy) * n + xx) * CIN + 8 * c8);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) xT[(8 * c8 + i) * PB + j] = val[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TW; ++i) {
            const int t = w + 4 * i;
            if (t < T) {  // uniform per wavefront
                const int mt = t / KT, kt = t % KT;
#pragma unroll
                for (int h = 0; h < PB / 32; ++h) {
                    const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&dzT[(mt * 16 + nl) * PB + 32 * h + 8 * g]);
                    const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&xT[(kt * 16 + nl) * PB + 32 * h + 8 * g]);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // the staged blocks are rewritten next
    }
    // D[4g+i][nl] = dW[mt*16 + 4g + i][tap*CIN + kt*16 + nl]
    float *out = partial + (size_t)s * COUT * Kp;
#pragma unroll
    for (int i = 0; i < TW; ++i) {
        const int t = w + 4 * i;
        if (t < T) {
            const int mt = t / KT, kt = t % KT;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(size_t)(mt * 16 + 4 * g + r) * Kp + tap * CIN + kt * 16 + nl] = acc[i][r];
        }
    }
    if (tap == 26)
        for (int e = threadIdx.x; e < COUT * (Kp - 27 * CIN); e += 256) {
            const int co = e / (Kp - 27 * CIN), k = 27 * CIN + e % (Kp - 27 * CIN);
            out[(size_t)co * Kp + k] = 0.f;
        }
}

// Weight gradient for every tap at once, for grids whose x rows tile 64-point
// blocks (n in 16, 32, 64): workgroup s computes partial[s][co][tap*CIN + ci]
// over its slab.  Per block of 64 points (R = 64/n whole x rows of one plane)
// it stages dZ once, transposed [co][point], and the 3 x (R+2) input rows
// around the block once, transposed [ci][x] with a zero halo column at each
// end (x = -1 and x = n): the tap-shifted input of every tap is then a window
// of those rows, so nothing is staged per tap.  Wavefront w owns the taps of
// WROWS (dz, dy) rows; a lane's B fragment (8 consecutive points of one input
// channel, shifted by dx) comes from two aligned 16-byte LDS reads: dx = -1 is
// the first, dx = +1 the dwords 1..4, dx = 0 a 2-byte byte-align of the pair.
// Input channels below 16 are padded with zero rows (the first layer's 8).
// KS > 1 splits the input channels over KS workgroups per slab (blockIdx.y):
// each stages and computes CIN / KS of them, which keeps the accumulators of
// the wide layers in registers.
template <int CIN, int COUT, int WROWS, int KS = 1>
__global__ void __launch_bounds__(64 * 9 / WROWS) conv3_wgrad_rows_kernel(const bf16 *__restrict__ Xg,
                                                                        const bf16 *__restrict__ dZ, int Kp,
                                                                        float *__restrict__ partial, int n, int npts,
                                                                        int slab) {
    constexpr int CINX = CIN;            // channels per point of the input tensor
    constexpr int CW = CIN / KS;         // input channels of this workgroup
    constexpr int CINP = CW < 16 ? 16 : CW;
    constexpr int MT = COUT / 16, KT = CINP / 16, NWAVE = 9 / WROWS, NTH = 64 * NWAVE;
    static_assert(CIN % KS == 0 && (KS == 1 || CW % 16 == 0), "channel split");
    const int ci0 = blockIdx.y * CW;
    const bf16 *X = Xg + ci0;
    constexpr int DS = 72;  // dZ row stride (64 points + 8: spreads the 16 rows of a fragment over the banks)
    static_assert(CINP % 16 == 0 && COUT % 16 == 0 && CW % 8 == 0 && 9 % WROWS == 0, "shape");
    extern __shared__ __attribute__((aligned(16))) bf16 wg_lds[];
    const int R = 64 / n, W = n + 8, ROWS = 3 * (R + 2);
    bf16 *dzT = wg_lds;                 // [COUT][DS]
    bf16 *xT = wg_lds + COUT * DS;      // [ROWS][CINP][W], x + 1 stored at column x + 1
    const int s = blockIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nl = lane & 15, g = lane >> 4;
    const bf16x8 zero8 = {(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f,
                          (bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
    if (CW < CINP) {  // padded input channels: zero once, never overwritten
        for (int v = threadIdx.x; v < ROWS * (CINP - CW) * W; v += NTH) {
            const int c = CW + (v / W) % (CINP - CW), r = v / (W * (CINP - CW));
            xT[((size_t)r * CINP + c) * W + v % W] = (bf16)0.0f;
        }
    }
    f32x4 acc[WROWS][3][MT][KT];
#pragma unroll
    for (int a = 0; a < WROWS; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int k = 0; k < KT; ++k) acc[a][d][m][k] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int n2 = n * n, n3 = n2 * n;
    const int pbeg = s * slab, pend = min(npts, pbeg + slab);
    for (int pb = pbeg; pb < pend; pb += 64) {
        const int b = pb / n3, rem = pb - b * n3;
        const int z = rem / n2, y0 = (rem / n) % n;
        // stage dZ of the 64 points, transposed: two points per thread, so each LDS store
        // writes one channel's pair of neighbouring points (4 bytes)
        for (int v = threadIdx.x; v < 32 * (COUT / 8); v += NTH) {
            const int j = 2 * (v / (COUT / 8)), c8 = v % (COUT / 8);
            const u16x8 a0 = __builtin_bit_cast(u16x8, *reinterpret_cast<const bf16x8 *>(dZ + (size_t)(pb + j) * COUT + 8 * c8));
            const u16x8 a1 =
                __builtin_bit_cast(u16x8, *reinterpret_cast<const bf16x8 *>(dZ + (size_t)(pb + j + 1) * COUT + 8 * c8));
#pragma unroll
            for (int i = 0; i < 8; ++i)
                *reinterpret_cast<unsigned *>(&dzT[(8 * c8 + i) * DS + j]) = (unsigned)a0[i] | ((unsigned)a1[i] << 16);
        }
        // stage the 3 x (R+2) input rows around the block, transposed, x = -1 .. n (n + 2 columns,
        // two per thread as above)
        for (int v = threadIdx.x; v < ROWS * (n / 2 + 1) * (CW / 8); v += NTH) {
            const int c8 = v % (CW / 8);
            const int xj = 2 * ((v / (CW / 8)) % (n / 2 + 1));
            const int r = v / ((CW / 8) * (n / 2 + 1));
            const int zz = z + r / (R + 2) - 1, yy = y0 + r % (R + 2) - 1;
            u16x8 a[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int xx = xj + e - 1;
                bf16x8 val = zero8;
                if (xx >= 0 && xx < n && yy >= 0 && yy < n && zz >= 0 && zz < n)
                    val = *reinterpret_cast<const bf16x8 *>(X + ((((size_t)b * n + zz) * n + yy) * n + xx) * CINX + 8 * c8);
                a[e] = __builtin_bit_cast(u16x8, val);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
                *reinterpret_cast<unsigned *>(&xT[((size_t)r * CINP + 8 * c8 + i) * W + xj]) =
                    (unsigned)a[0][i] | ((unsigned)a[1][i] << 16);
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // two K steps of 32 points
            const int q = 32 * h + 8 * g;  // this lane's first point in the block
            const int rq = q / n, xg = q % n;
            bf16x8 av[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const bf16x8 *>(&dzT[(m * 16 + nl) * DS + q]);
#pragma unroll
            for (int a = 0; a < WROWS; ++a) {
                const int rw = w * WROWS + a;  // (dz, dy) row index 0..8
                const int dz = rw / 3, dy = rw % 3;  // offsets + 1
                const int r = dz * (R + 2) + rq + dy;
#pragma unroll
                for (int k = 0; k < KT; ++k) {
                    const bf16 *src = &xT[((size_t)r * CINP + k * 16 + nl) * W + xg];
                    const bf16x8 lo = *reinterpret_cast<const bf16x8 *>(src);
                    const bf16x8 hi = *reinterpret_cast<const bf16x8 *>(src + 8);
                    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
                    const u32x4 l4 = __builtin_bit_cast(u32x4, lo), h4 = __builtin_bit_cast(u32x4, hi);
                    const u32x4 mid = {__builtin_amdgcn_alignbyte(l4[1], l4[0], 2),
                                       __builtin_amdgcn_alignbyte(l4[2], l4[1], 2),
                                       __builtin_amdgcn_alignbyte(l4[3], l4[2], 2),
                                       __builtin_amdgcn_alignbyte(h4[0], l4[3], 2)};
                    const u32x4 up = {l4[1], l4[2], l4[3], h4[0]};
                    const bf16x8 bv[3] = {lo, __builtin_bit_cast(bf16x8, mid), __builtin_bit_cast(bf16x8, up)};
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int m = 0; m < MT; ++m)
                            acc[a][d][m][k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[d], acc[a][d][m][k], 0,
                                                                                      0, 0);
                }
            }
        }
        __syncthreads();  // the staged rows are rewritten next
    }
    // D[4g+i][nl] = dW[m*16 + 4g + i][tap*CIN + k*16 + nl], tap = 9 dz + 3 dy + dx (offsets + 1)
    float *out = partial + (size_t)s * COUT * Kp;
#pragma unroll
    for (int a = 0; a < WROWS; ++a) {
        const int rw = w * WROWS + a;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int tap = 3 * rw + d;
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int k = 0; k < KT; ++k) {
                    const int ci = k * 16 + nl;
                    if (ci < CW)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            out[(size_t)(m * 16 + 4 * g + i) * Kp + tap * CIN + ci0 + ci] = acc[a][d][m][k][i];
                }
        }
    }
    if (blockIdx.y == 0)
    for (int e = threadIdx.x; e < COUT * (Kp - 27 * CIN); e += NTH) {  // the zero-padded K tail
        const int co = e / (Kp - 27 * CIN), k = 27 * CIN + e % (Kp - 27 * CIN);
        out[(size_t)co * Kp + k] = 0.f;
    }
}

// out[i] = sum_s partial[s][i] (fixed order s = 0, 1, ...), i < count.  Eight
// loads are issued before their eight in-order adds, so the memory latency is
// paid once per eight slabs instead of once per slab.
__global__ void reduce_slabs_kernel(const float *__restrict__ partial, int nslab, size_t count,
                                    float *__restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    float a = 0.f;
    int s = 0;
    for (; s + 8 <= nslab; s += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = partial[(size_t)(s + j) * count + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) a += v[j];
    }
    for (; s < nslab; ++s) a += partial[(size_t)s * count + i];
    out[i] = a;
}

// First pass of a two-pass slab reduction: part p of P sums slabs
// [p*nslab/P, (p+1)*nslab/P) in order into tmp[p][i], so P times more threads
// share the reads; reduce_slabs_kernel then adds the P parts in order (a fixed
// tree: the same bits on every call).
__global__ void reduce_slab_parts_kernel(const float *__restrict__ partial, int nslab, size_t count, int nparts,
                                         float *__restrict__ tmp) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const int part = blockIdx.y;
    if (i >= count) return;
    const int s0 = (int)((long long)nslab * part / nparts), s1 = (int)((long long)nslab * (part + 1) / nparts);
    float a = 0.f;
    int s = s0;
    for (; s + 8 <= s1; s += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = partial[(size_t)(s + j) * count + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) a += v[j];
    }
    for (; s < s1; ++s) a += partial[(size_t)s * count + i];
    tmp[(size_t)part * count + i] = a;
}

// out[r] = sum of row r of part ([rows][nslab], row-major), one 1024-thread
// workgroup per row: thread t adds s = t, t + 1024, ... in order (eight loads in
// flight), then a fixed pairwise tree over the threads, so every call gives the
// same bits.  For the few-value sums (bias and head gradients) whose slab
// counts run to tens of thousands.
__global__ void __launch_bounds__(1024) reduce_rows_kernel(const float *__restrict__ part, int nslab,
                                                           float *__restrict__ out) {
    __shared__ float sh[1024];
    const float *row = part + (size_t)blockIdx.x * nslab;
    const int t = threadIdx.x;
    float a = 0.f;
    int s = t;
    for (; s + 7 * 1024 < nslab; s += 8 * 1024) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = row[s + j * 1024];
#pragma unroll
        for (int j = 0; j < 8; ++j) a += v[j];
    }
    for (; s < nslab; s += 1024) a += row[s];
    sh[t] = a;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) out[blockIdx.x] = sh[0];
}

// per-slab channel sums of dZ (bias gradient), partial[c][slab] (rows for reduce_rows_kernel).  256 threads
// per slab: thread t < R*C (R = 256 / C rows at a time) owns channel t % C and
// rows t / C, t / C + R, ... of the slab, so each pass reads 256 consecutive
// values; the R row-partials of a channel are then added in row order.
__global__ void __launch_bounds__(256) chan_sum_kernel(const bf16 *__restrict__ dZ, int C, int npts, int slab,
                                                       float *__restrict__ partial) {
    __shared__ float red[256];
    const int s = blockIdx.x, t = threadIdx.x;
    const int R = 256 / C;  // C <= 256
    const int pbeg = s * slab, pend = min(npts, pbeg + slab);
    float a = 0.f;
    if (t < R * C) {
        const int c = t % C;
        int p = pbeg + t / C;
        for (; p + 7 * R < pend; p += 8 * R) {  // eight loads in flight, adds in row order
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = bf2f(dZ[(size_t)(p + j * R) * C + c]);
#pragma unroll
            for (int j = 0; j < 8; ++j) a += v[j];
        }
        for (; p < pend; p += R) a += bf2f(dZ[(size_t)p * C + c]);
    }
    red[t] = a;
    __syncthreads();
    if (t < C) {
        float v = 0.f;
        for (int r = 0; r < R; ++r) v += red[r * C + t];
        partial[(size_t)t * gridDim.x + s] = v;
    }
}

// --------------------------------------------------------- elementwise ----
// The elementwise kernels below move 8 channels (one 16-byte vector) per
// thread; every channel count and offset of the network is a multiple of 8
// and points number fewer than 2^27, so 32-bit index math suffices.
__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ void st8(bf16 *p, bf16x8 v) { *reinterpret_cast<bf16x8 *>(p) = v; }

// 2x2x2 average pooling: n -> n/2
__global__ void pool_fwd_kernel(const bf16 *__restrict__ X, bf16 *__restrict__ Y, int C, int n, int nb) {
    const unsigned m = n / 2, V = C / 8;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)nb * m * m * m * V) return;
    const unsigned cv = i % V, q = i / V;
    const unsigned x = q % m, y = (q / m) % m, z = (q / (m * m)) % m, b = q / (m * m * m);
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const bf16x8 v = ld8(X + ((((size_t)b * n + 2 * z + dz) * n + 2 * y + dy) * n + 2 * x + dx) * C + 8 * cv);
#pragma unroll
                for (int j = 0; j < 8; ++j) a[j] += bf2f(v[j]);
            }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = f2bf(a[j] * 0.125f);
    st8(Y + (size_t)q * C + 8 * cv, o);
}

// gradient of 2x2x2 average pooling, then (skip != null) plus the skip connection's gradient,
// channels [skip_off, skip_off + C) of a buffer with skip_ct channels, then (mask != null) the
// ReLU backward of the layer that produced the pooled tensor: zero where mask ([points][C]) is
// not positive.  Each step rounds to BF16 as separate kernels would.
__global__ void pool_bwd_kernel(const bf16 *__restrict__ dY, bf16 *__restrict__ dX, int C, int n, int nb,
                                const bf16 *__restrict__ skip, int skip_ct, int skip_off,
                                const bf16 *__restrict__ mask) {
    const unsigned V = C / 8, un = n, m = n / 2;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)nb * un * un * un * V) return;
    const unsigned cv = i % V, q = i / V;
    const unsigned x = q % un, y = (q / un) % un, z = (q / (un * un)) % un, b = q / (un * un * un);
    const bf16x8 v = ld8(dY + ((((size_t)b * m + z / 2) * m + y / 2) * m + x / 2) * C + 8 * cv);
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = f2bf(bf2f(v[j]) * 0.125f);
    if (skip) {
        const bf16x8 sk = ld8(skip + (size_t)q * skip_ct + skip_off + 8 * cv);
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = f2bf(bf2f(o[j]) + bf2f(sk[j]));
    }
    if (mask) {
        const bf16x8 mk = ld8(mask + (size_t)q * C + 8 * cv);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (!(bf2f(mk[j]) > 0.f)) o[j] = (bf16)0.0f;
    }
    st8(dX + (size_t)q * C + 8 * cv, o);
}

// nearest upsampling m -> 2m, written into channels [off, off+C) of a buffer with Ct channels
__global__ void up_fwd_kernel(const bf16 *__restrict__ X, bf16 *__restrict__ Y, int C, int Ct, int off, int m,
                              int nb) {
    const unsigned n = 2 * m, V = C / 8, um = m;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)nb * n * n * n * V) return;
    const unsigned cv = i % V, q = i / V;
    const unsigned x = q % n, y = (q / n) % n, z = (q / (n * n)) % n, b = q / (n * n * n);
    st8(Y + (size_t)q * Ct + off + 8 * cv, ld8(X + ((((size_t)b * um + z / 2) * um + y / 2) * um + x / 2) * C + 8 * cv));
}

// gradient of nearest upsampling: sum of the 8 children (from channels [off, off+C) of dY), then
// (mask != null) the ReLU backward of the layer that produced the upsampled tensor
__global__ void up_bwd_kernel(const bf16 *__restrict__ dY, bf16 *__restrict__ dX, int C, int Ct, int off, int m,
                              int nb, const bf16 *__restrict__ mask) {
    const unsigned n = 2 * m, V = C / 8, um = m;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)nb * um * um * um * V) return;
    const unsigned cv = i % V, q = i / V;
    const unsigned x = q % um, y = (q / um) % um, z = (q / (um * um)) % um, b = q / (um * um * um);
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const bf16x8 v =
                    ld8(dY + ((((size_t)b * n + 2 * z + dz) * n + 2 * y + dy) * n + 2 * x + dx) * Ct + off + 8 * cv);
#pragma unroll
                for (int j = 0; j < 8; ++j) a[j] += bf2f(v[j]);
            }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = f2bf(a[j]);
    if (mask) {
        const bf16x8 mk = ld8(mask + (size_t)q * C + 8 * cv);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (!(bf2f(mk[j]) > 0.f)) o[j] = (bf16)0.0f;
    }
    st8(dX + (size_t)q * C + 8 * cv, o);
}

// copy C channels between buffers with different channel counts / offsets
__global__ void chan_copy_kernel(const bf16 *__restrict__ src, int Cs, int offs, bf16 *__restrict__ dst, int Cd,
                                 int offd, int C, size_t npts) {
    const unsigned V = C / 8;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned)npts * V) return;
    const unsigned p = i / V, cv = i % V;
    st8(dst + (size_t)p * Cd + offd + 8 * cv, ld8(src + (size_t)p * Cs + offs + 8 * cv));
}

// input packing: host-order float planes [b][2][n^3] -> NDHWC BF16 with 8 channels (2 used)
__global__ void pack_input_kernel(const float *__restrict__ in, bf16 *__restrict__ X, int n3, int nb) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)nb * n3) return;
    const size_t b = i / n3, q = i % n3;
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (bf16)0.0f;
    v[0] = f2bf(in[(b * 2 + 0) * n3 + q]);
    v[1] = f2bf(in[(b * 2 + 1) * n3 + q]);
    *reinterpret_cast<bf16x8 *>(X + i * 8) = v;
}

// 1x1 head 16 -> 1: out[p] = sum_c w[c] X[p][c] + w[16] (FP32 out)
__global__ void head_fwd_kernel(const bf16 *__restrict__ X, const float *__restrict__ w, float *__restrict__ out,
                                size_t npts) {
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= npts) return;
    float a = w[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) a += w[c] * bf2f(X[p * 16 + c]);
    out[p] = a;
}

// loss = mean((pred - target)^2); dpred = 2 (pred - target) / N  (partials per block)
__global__ void mse_kernel(const float *__restrict__ pred, const float *__restrict__ target, float *__restrict__ dpred,
                           double *__restrict__ lpart, size_t npts) {
    __shared__ double sh[256];
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    double l = 0.0;
    if (p < npts) {
        const float d = pred[p] - target[p];
        dpred[p] = 2.0f * d / (float)npts;
        l = (double)d * d;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// head backward: dX[p][c] = dpred[p] w[c] (BF16), zero where X (the ReLU output the head reads)
// is not positive (the ReLU backward of the last conv); per-block partial grads of w (16) and
// b (1), gpart[c][block] (rows for reduce_rows_kernel)
__global__ void head_bwd_kernel(const bf16 *__restrict__ X, const float *__restrict__ w,
                                const float *__restrict__ dpred, bf16 *__restrict__ dX, float *__restrict__ gpart,
                                size_t npts) {
    __shared__ float sh[17][256];
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    float gw[17];
#pragma unroll
    for (int c = 0; c < 17; ++c) gw[c] = 0.f;
    if (p < npts) {
        const float d = dpred[p];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bf16x8 x = ld8(X + p * 16 + 8 * h);
            bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                o[j] = bf2f(x[j]) > 0.f ? f2bf(d * w[8 * h + j]) : (bf16)0.0f;
                gw[8 * h + j] = d * bf2f(x[j]);
            }
            st8(dX + p * 16 + 8 * h, o);
        }
        gw[16] = d;
    }
#pragma unroll
    for (int c = 0; c < 17; ++c) sh[c][threadIdx.x] = gw[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int c = 0; c < 17; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 17) gpart[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// Adam on FP32 master parameters
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, size_t count, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master conv weights [Cout][Cin_real][27] -> forward layout [Cout][Kp] (k = tap*Cin + ci)
// and backward-data layout [Cin][Kp_t] (k = tap'*Cout + co, tap' = 26 - tap), BF16
__global__ void pack_weights_kernel(const float *__restrict__ w, int Cout, int Cin_real, int Cin, int Kp,
                                    int Kp_t, bf16 *__restrict__ Wf, bf16 *__restrict__ Wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = Cout * Cin * 27;
    if (i >= total) return;
    const int tap = i % 27, ci = (i / 27) % Cin, co = i / (27 * Cin);
    const float v = ci < Cin_real ? w[((size_t)co * Cin_real + ci) * 27 + tap] : 0.f;
    Wf[(size_t)co * Kp + tap * Cin + ci] = f2bf(v);
    if (Wt) Wt[(size_t)ci * Kp_t + (26 - tap) * Cout + co] = f2bf(v);
}

// weight-gradient layout [Cout][Kp] (k = tap*Cin + ci) -> master order [Cout][Cin_real][27]
__global__ void unpack_wgrad_kernel(const float *__restrict__ gk, int Cout, int Cin_real, int Cin, int Kp,
                                    float *__restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Cout * Cin_real * 27) return;
    const int tap = i % 27, ci = (i / 27) % Cin_real, co = i / (27 * Cin_real);
    g[i] = gk[(size_t)co * Kp + tap * Cin + ci];
}

}  // namespace unet
}  // namespace cfm

using namespace cfm;
using namespace cfm::unet;

namespace {

// Layer table: 10 3x3x3 convolutions (+ReLU) and a 1x1 head.
struct ConvSpec {
    int cin_real, cin, cout, level;
};
constexpr int kNConv = 10;
constexpr ConvSpec kConv[kNConv] = {
    {2, 8, 16, 0},   {16, 16, 16, 0},  // encoder level 1 (input padded to 8 channels)
    {16, 16, 32, 1}, {32, 32, 32, 1},  // encoder level 2
    {32, 32, 64, 2}, {64, 64, 64, 2},  // bottleneck
    {96, 96, 32, 1}, {32, 32, 32, 1},  // decoder level 2 (64 upsampled + 32 skip)
    {48, 48, 16, 0}, {16, 16, 16, 0},  // decoder level 1 (32 upsampled + 16 skip)
};
constexpr int kHeadIn = 16;

inline int pad32(int k) { return (k + 31) / 32 * 32; }

constexpr int kReduceParts = 16;  // parts of the first pass of the slab reductions

size_t param_count() {
    size_t c = 0;
    for (const auto &s : kConv) c += (size_t)s.cout * s.cin_real * 27 + s.cout;
    return c + kHeadIn + 1;
}

inline dim3 blocks_for(size_t count, int threads = 256) { return dim3((unsigned)((count + threads - 1) / threads)); }

}  // namespace

struct cfm_unet {
    int n = 0, batch = 0;
    size_t P[3] = {0, 0, 0};  // points per level (batch included)
    hipStream_t stream = nullptr;
    // parameters
    float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    size_t nparam = 0;
    size_t woff[kNConv], boff[kNConv], head_off = 0;
    bf16 *Wf[kNConv] = {}, *Wt[kNConv] = {};
    int Kp[kNConv], Kpt[kNConv];
    float *gk = nullptr;  // weight gradient in kernel layout (largest layer)
    float *rtmp = nullptr;  // two-pass slab reduction: [kReduceParts][largest layer]
    float *partial = nullptr;
    size_t partial_cap = 0;
    int slab = 4096;
    // activations
    bf16 *X0 = nullptr, *A0 = nullptr, *S1 = nullptr, *Q1 = nullptr, *A2 = nullptr, *S2 = nullptr, *Q2 = nullptr,
         *A4 = nullptr, *A5 = nullptr, *U2 = nullptr, *A6 = nullptr, *A7 = nullptr, *U1 = nullptr, *A8 = nullptr,
         *A9 = nullptr;
    float *pred = nullptr, *target = nullptr, *dpred = nullptr, *in_f = nullptr;
    double *lpart = nullptr;
    float *hpart = nullptr;
    // gradients of activations (BF16), sized for the largest of each level
    bf16 *g1a = nullptr, *g1b = nullptr, *g1c = nullptr;  // level 1, up to 48 channels
    bf16 *g2a = nullptr, *g2b = nullptr, *g2c = nullptr;  // level 2, up to 96 channels
    bf16 *g3a = nullptr, *g3b = nullptr;                                   // level 3, 64 channels
    long long step = 0;
    std::vector<void *> allocs;
};

static cfm_status unet_alloc(cfm_unet *u, void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) return CFM_ENOMEM;
    u->allocs.push_back(*p);
    return CFM_OK;
}

extern "C" void cfm_unet_destroy(cfm_unet *u) {
    if (!u) return;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (void *p : u->allocs) (void)hipFree(p);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

extern "C" size_t cfm_unet_param_count(void) { return param_count(); }

extern "C" cfm_status cfm_unet_create(int n, int batch, cfm_unet **out) {
    if (!out) return CFM_EINVAL;
    *out = nullptr;
    // batch * n^3 <= 2^27 keeps the weight-gradient slab count (points / 4096) within a grid's y extent
    if (n < 4 || n % 4 != 0 || batch < 1 || (long long)batch * n * n * n > (1LL << 27)) return CFM_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CFM_EDEVICE;
    auto *u = new (std::nothrow) cfm_unet();
    if (!u) return CFM_ENOMEM;
    auto fail = [&](cfm_status s) {
        cfm_unet_destroy(u);
        return s;
    };
    u->n = n;
    u->batch = batch;
    for (int l = 0; l < 3; ++l) {
        const size_t m = (size_t)(n >> l);
        u->P[l] = (size_t)batch * m * m * m;
    }
    if (hipStreamCreate(&u->stream) != hipSuccess) return fail(CFM_EDEVICE);
    u->nparam = param_count();
    size_t off = 0;
    for (int i = 0; i < kNConv; ++i) {
        u->woff[i] = off;
        off += (size_t)kConv[i].cout * kConv[i].cin_real * 27;
        u->boff[i] = off;
        off += kConv[i].cout;
        u->Kp[i] = pad32(27 * kConv[i].cin);
        u->Kpt[i] = pad32(27 * kConv[i].cout);
    }
    u->head_off = off;
    cfm_status st = CFM_OK;
    auto A = [&](auto **p, size_t bytes) {
        if (st == CFM_OK) st = unet_alloc(u, (void **)p, bytes);
    };
    A(&u->master, 4 * u->nparam);
    A(&u->grad, 4 * u->nparam);
    A(&u->adam_m, 4 * u->nparam);
    A(&u->adam_v, 4 * u->nparam);
    size_t max_wk = 0, max_partial = 0;
    for (int i = 0; i < kNConv; ++i) {
        A(&u->Wf[i], 2 * (size_t)kConv[i].cout * u->Kp[i]);
        if (i > 0) A(&u->Wt[i], 2 * (size_t)kConv[i].cin * u->Kpt[i]);
        max_wk = std::max(max_wk, (size_t)kConv[i].cout * u->Kp[i]);
        const size_t nslab = (u->P[kConv[i].level] + u->slab - 1) / u->slab;
        max_partial = std::max(max_partial, nslab * kConv[i].cout * u->Kp[i]);
    }
    A(&u->gk, 4 * max_wk);
    A(&u->rtmp, 4 * kReduceParts * max_wk);
    A(&u->partial, 4 * max_partial);
    u->partial_cap = max_partial;
    const size_t P1 = u->P[0], P2 = u->P[1], P3 = u->P[2];
    A(&u->X0, 2 * P1 * 8);
    A(&u->A0, 2 * P1 * 16);
    A(&u->S1, 2 * P1 * 16);
    A(&u->Q1, 2 * P2 * 16);
    A(&u->A2, 2 * P2 * 32);
    A(&u->S2, 2 * P2 * 32);
    A(&u->Q2, 2 * P3 * 32);
    A(&u->A4, 2 * P3 * 64);
    A(&u->A5, 2 * P3 * 64);
    A(&u->U2, 2 * P2 * 96);
    A(&u->A6, 2 * P2 * 32);
    A(&u->A7, 2 * P2 * 32);
    A(&u->U1, 2 * P1 * 48);
    A(&u->A8, 2 * P1 * 16);
    A(&u->A9, 2 * P1 * 16);
    A(&u->pred, 4 * P1);
    A(&u->target, 4 * P1);
    A(&u->dpred, 4 * P1);
    A(&u->in_f, 4 * P1 * 2);
    A(&u->lpart, 8 * ((P1 + 255) / 256));
    A(&u->hpart, 4 * 17 * ((P1 + 255) / 256));
    for (bf16 **g : {&u->g1a, &u->g1b, &u->g1c}) A(g, 2 * P1 * 48);
    for (bf16 **g : {&u->g2a, &u->g2b, &u->g2c}) A(g, 2 * P2 * 96);
    for (bf16 **g : {&u->g3a, &u->g3b}) A(g, 2 * P3 * 64);
    if (st != CFM_OK) return fail(st);
    if (hipMemset(u->master, 0, 4 * u->nparam) != hipSuccess || hipMemset(u->adam_m, 0, 4 * u->nparam) != hipSuccess ||
        hipMemset(u->adam_v, 0, 4 * u->nparam) != hipSuccess)
        return fail(CFM_EDEVICE);
    *out = u;
    return CFM_OK;
}

static cfm_status pack_all_weights(cfm_unet *u) {
    for (int i = 0; i < kNConv; ++i) {
        const ConvSpec &s = kConv[i];
        CFM_HIP(hipMemsetAsync(u->Wf[i], 0, 2 * (size_t)s.cout * u->Kp[i], u->stream));
        if (u->Wt[i]) CFM_HIP(hipMemsetAsync(u->Wt[i], 0, 2 * (size_t)s.cin * u->Kpt[i], u->stream));
        const int total = s.cout * s.cin * 27;
        hipLaunchKernelGGL(pack_weights_kernel, blocks_for(total), dim3(256), 0, u->stream, u->master + u->woff[i],
                           s.cout, s.cin_real, s.cin, u->Kp[i], u->Kpt[i], u->Wf[i], u->Wt[i]);
        CFM_HIP(hipGetLastError());
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_unet_set_params(cfm_unet *u, const float *params) {
    if (!u || !params) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(u->master, params, 4 * u->nparam, hipMemcpyHostToDevice));
    CFM_HIP(hipMemset(u->adam_m, 0, 4 * u->nparam));
    CFM_HIP(hipMemset(u->adam_v, 0, 4 * u->nparam));
    u->step = 0;
    cfm_status st = pack_all_weights(u);
    if (st != CFM_OK) return st;
    CFM_HIP(hipStreamSynchronize(u->stream));
    return CFM_OK;
}

extern "C" cfm_status cfm_unet_get_params(cfm_unet *u, float *params) {
    if (!u || !params) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(params, u->master, 4 * u->nparam, hipMemcpyDeviceToHost));
    return CFM_OK;
}

template <int MT, int CIN>
static bool launch_conv_lds(int n, int batch, hipStream_t st, const bf16 *X, const bf16 *W, int kp,
                            const float *bias, bf16 *Y, int relu, const bf16 *mask) {
    if (kp != (27 * CIN + 31) / 32 * 32) return false;
    // tile 16 x TY x TZ points, the staged block at most 80 KiB of LDS (two workgroups per CU): up
    // to 48 input channels 16 x 8 x 2 (four rows per wavefront), 64 channels 16 x 4 x 2, 96
    // channels 16 x 4 x 1
    constexpr int TY = CIN <= 48 ? 8 : 4;
    constexpr int TZ = CIN <= 64 ? 2 : 1;
    static_assert((TZ + 2) * (TY + 2) * 18 * CIN * 2 <= 80 * 1024, "staged block above 80 KiB");
    const int nbx = (n + 15) / 16, nby = (n + TY - 1) / TY, nbz = (n + TZ - 1) / TZ;
    const dim3 grid((unsigned)(nbx * nby * nbz * batch));
    hipLaunchKernelGGL((conv3_fwd_lds_kernel<MT, CIN, TY, TZ>), grid, dim3(256), 0, st, X, W, kp, bias, Y, n, nbx, nby,
                       nbz, relu, mask);
    return true;
}

#ifndef CFM_UNET_XACC
#define CFM_UNET_XACC 1
#endif

// Rows of NXT 16-point tiles: eight (y, z) rows (4 x 2), two per wavefront, while the
// accumulators (3 MT x 2 x NXT tiles) stay within 24 tiles, else 2 x 2 rows, one per
// wavefront; eight wavefronts (one row each) when the eight-row block is too large for two
// workgroups per CU.
template <int MT, int CIN, int NXT>
static void launch_conv_xacc(int n, int batch, hipStream_t st, const bf16 *X, const bf16 *W, int kp,
                             const float *bias, bf16 *Y, int relu, const bf16 *mask) {
    constexpr bool two = MT * NXT <= 4;
    constexpr int lds8 = NXT * 16 * 6 * 4 * CIN * 2;
    constexpr int NW = two && lds8 > 80 * 1024 ? 8 : 4;
    constexpr int TY = two ? 4 : 2, TZ = 2;
    static_assert(NXT * 16 * (TY + 2) * (TZ + 2) * CIN * 2 <= 144 * 1024, "staged block above 144 KiB");
    const int nby = (n + TY - 1) / TY, nbz = (n + TZ - 1) / TZ;
    hipLaunchKernelGGL((conv3_fwd_xacc_kernel<MT, CIN, TY, TZ, NXT, NW, 2>), dim3((unsigned)(nby * nbz * batch)),
                       dim3(64 * NW), 0, st, X, W, kp, bias, Y, n, nby, nbz, relu, mask);
}

// The x-accumulator kernel where it measured faster than the per-tap one: the wide inputs of
// the decoder, 48 -> 16 channels (rows of 64 points at 32 x 64^3: 1.06 vs 1.17 ms) and 96 -> 32
// (rows of 32: 0.59 vs 0.93 ms).  With 8 to 32 input channels the per-tap kernel's smaller
// staged block keeps 3 to 7 workgroups per CU against 1 to 3 and won (16 -> 16: 0.33 vs 0.41
// ms), and the 16 -> 48 and 32 -> 96 data gradients need 3 x 3 and 3 x 6 accumulator tiles per
// row: one wavefront per SIMD, twice the time.
static bool conv_xacc(int cin, int mt, int n, int batch, hipStream_t st, const bf16 *X, const bf16 *W, int kp,
                      const float *bias, bf16 *Y, int relu, const bf16 *mask) {
    if (!CFM_UNET_XACC || n > 64 || kp < 27 * cin) return false;
    const int nxt = n <= 16 ? 1 : n <= 32 ? 2 : 4;
#define CFM_CONV_XACC(C, M, T)                                               \
    if (cin == C && mt == M && nxt == T) {                                   \
        launch_conv_xacc<M, C, T>(n, batch, st, X, W, kp, bias, Y, relu, mask); \
        return true;                                                         \
    }
    CFM_CONV_XACC(48, 1, 4)
    CFM_CONV_XACC(48, 1, 2)
    CFM_CONV_XACC(48, 1, 1)
    CFM_CONV_XACC(96, 2, 2)
    CFM_CONV_XACC(96, 2, 1)
#undef CFM_CONV_XACC
    return false;
}

// the (input channels, output channel tiles) pairs of the network's forward and data-gradient convs
static bool conv_lds(int cin, int mt, int n, int batch, hipStream_t st, const bf16 *X, const bf16 *W, int kp,
                     const float *bias, bf16 *Y, int relu, const bf16 *mask) {
#define CFM_CONV_LDS(C, M)                                                  \
    if (cin == C && mt == M) {                                              \
        return launch_conv_lds<M, C>(n, batch, st, X, W, kp, bias, Y, relu, mask); \
    }
    CFM_CONV_LDS(8, 1)
    CFM_CONV_LDS(16, 1)
    CFM_CONV_LDS(16, 2)
    CFM_CONV_LDS(16, 3)
    CFM_CONV_LDS(32, 1)
    CFM_CONV_LDS(32, 2)
    CFM_CONV_LDS(32, 4)
    CFM_CONV_LDS(32, 6)
    CFM_CONV_LDS(48, 1)
    CFM_CONV_LDS(64, 2)
    CFM_CONV_LDS(64, 4)
    CFM_CONV_LDS(96, 2)
#undef CFM_CONV_LDS
    return false;
}

// Y = conv(X) for the (input channels, output channels) pairs of the network
// mask (may be null): the data gradient of a ReLU layer, zero where mask (the layer's forward
// output, [points][cout]) is not positive
static cfm_status conv_fwd(cfm_unet *u, const bf16 *X, int cin, const bf16 *W, int kp, const float *bias, bf16 *Y,
                           int cout, int level, int relu, const bf16 *mask = nullptr) {
    const int n = u->n >> level;
    if (cout % 16 != 0 || !(conv_xacc(cin, cout / 16, n, u->batch, u->stream, X, W, kp, bias, Y, relu, mask) ||
                            conv_lds(cin, cout / 16, n, u->batch, u->stream, X, W, kp, bias, Y, relu, mask)))
        return CFM_FAIL(CFM_EINVAL, "unet: no convolution kernel for %d -> %d channels", cin, cout);
    CFM_HIP(hipGetLastError());
    return CFM_OK;
}

static cfm_status run_forward_dev(cfm_unet *u, const float *d_in) {
    const size_t P1 = u->P[0], P2 = u->P[1];
    const int n = u->n, n3 = n * n * n;
    hipLaunchKernelGGL(pack_input_kernel, blocks_for(P1), dim3(256), 0, u->stream, d_in, u->X0, n3, u->batch);
    CFM_HIP(hipGetLastError());
    const float *b[kNConv];
    for (int i = 0; i < kNConv; ++i) b[i] = u->master + u->boff[i];
    cfm_status st;
#define CONV(I, X, Y) \
    if ((st = conv_fwd(u, X, kConv[I].cin, u->Wf[I], u->Kp[I], b[I], Y, kConv[I].cout, kConv[I].level, 1)) != CFM_OK) return st
    CONV(0, u->X0, u->A0);
    CONV(1, u->A0, u->S1);
    hipLaunchKernelGGL(pool_fwd_kernel, blocks_for(P2 * 16 / 8), dim3(256), 0, u->stream, u->S1, u->Q1, 16, n, u->batch);
    CONV(2, u->Q1, u->A2);
    CONV(3, u->A2, u->S2);
    hipLaunchKernelGGL(pool_fwd_kernel, blocks_for(u->P[2] * 32 / 8), dim3(256), 0, u->stream, u->S2, u->Q2, 32, n / 2, u->batch);
    CONV(4, u->Q2, u->A4);
    CONV(5, u->A4, u->A5);
    hipLaunchKernelGGL(up_fwd_kernel, blocks_for(P2 * 64 / 8), dim3(256), 0, u->stream, u->A5, u->U2, 64, 96, 0, n / 4, u->batch);
    hipLaunchKernelGGL(chan_copy_kernel, blocks_for(P2 * 32 / 8), dim3(256), 0, u->stream, u->S2, 32, 0, u->U2, 96, 64, 32, P2);
    CONV(6, u->U2, u->A6);
    CONV(7, u->A6, u->A7);
    hipLaunchKernelGGL(up_fwd_kernel, blocks_for(P1 * 32 / 8), dim3(256), 0, u->stream, u->A7, u->U1, 32, 48, 0, n / 2, u->batch);
    hipLaunchKernelGGL(chan_copy_kernel, blocks_for(P1 * 16 / 8), dim3(256), 0, u->stream, u->S1, 16, 0, u->U1, 48, 32, 16, P1);
    CONV(8, u->U1, u->A8);
    CONV(9, u->A8, u->A9);
#undef CONV
    hipLaunchKernelGGL(head_fwd_kernel, blocks_for(P1), dim3(256), 0, u->stream, u->A9, u->master + u->head_off,
                       u->pred, P1);
    CFM_HIP(hipGetLastError());
    return CFM_OK;
}

static cfm_status run_forward(cfm_unet *u, const float *input_host) {
    CFM_HIP(hipMemcpyAsync(u->in_f, input_host, 4 * u->P[0] * 2, hipMemcpyHostToDevice, u->stream));
    return run_forward_dev(u, u->in_f);
}

namespace cfm {
UnetView unet_view(cfm_unet *u) { return UnetView{u->n, u->batch, u->stream, u->in_f}; }
cfm_status unet_forward_device(cfm_unet *u, const float *d_in, const float **d_pred) {
    cfm_status st = run_forward_dev(u, d_in);
    if (st == CFM_OK && d_pred) *d_pred = u->pred;
    return st;
}
}  // namespace cfm

extern "C" cfm_status cfm_unet_forward(cfm_unet *u, const float *input, float *output) {
    if (!u || !input || !output) return CFM_EINVAL;
    cfm_status st = run_forward(u, input);
    if (st != CFM_OK) return st;
    CFM_HIP(hipMemcpyAsync(output, u->pred, 4 * u->P[0], hipMemcpyDeviceToHost, u->stream));
    CFM_HIP(hipStreamSynchronize(u->stream));
    return CFM_OK;
}

// weight and bias gradients of conv I from its input X and output gradient dZ
// the LDS-staged weight gradient for the network's (input, output) channel pairs
// with 16-multiple input channels; the first layer (8 padded input channels)
// keeps the gather kernel
template <int CIN, int COUT, int WROWS, int KS = 1>
static void launch_wgrad_rows(int nslab, hipStream_t st, const bf16 *X, const bf16 *dZ, int kp, float *partial, int n,
                              int npts, int slab) {
    constexpr int CINP = CIN / KS < 16 ? 16 : CIN / KS;
    const int R = 64 / n;
    const size_t lds = sizeof(bf16) * ((size_t)COUT * 72 + (size_t)3 * (R + 2) * CINP * (n + 8));
    // (the largest staging, 48 input channels at n = 64, is just above 64 KiB)
    static const bool big_lds_ok =
        hipFuncSetAttribute((const void *)conv3_wgrad_rows_kernel<CIN, COUT, WROWS, KS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
    (void)big_lds_ok;
    hipLaunchKernelGGL((conv3_wgrad_rows_kernel<CIN, COUT, WROWS, KS>), dim3(nslab, KS), dim3(64 * 9 / WROWS), lds,
                       st, X, dZ, kp, partial, n, npts, slab);
}

// every tap in one workgroup (whole x rows per 64-point block): n = 16, 32, 64 and slabs of
// whole blocks; three (dz, dy) rows per wavefront up to 4 channel-tile pairs, else one
static bool wgrad_rows(int cin, int cout, int nslab, hipStream_t st, const bf16 *X, const bf16 *dZ, int kp,
                       float *partial, int n, int npts, int slab) {
    if ((n != 16 && n != 32 && n != 64) || slab % 64 != 0 || npts % 64 != 0) return false;
#define CFM_WGRAD_ROWS(C, O, WR, KS)                                                    \
    if (cin == C && cout == O) {                                                        \
        launch_wgrad_rows<C, O, WR, KS>(nslab, st, X, dZ, kp, partial, n, npts, slab);  \
        return true;                                                                    \
    }
    CFM_WGRAD_ROWS(8, 16, 3, 1)
    CFM_WGRAD_ROWS(16, 16, 3, 1)
    CFM_WGRAD_ROWS(16, 32, 3, 1)
    CFM_WGRAD_ROWS(32, 32, 3, 1)
    CFM_WGRAD_ROWS(48, 16, 3, 1)
    CFM_WGRAD_ROWS(32, 64, 1, 1)
    CFM_WGRAD_ROWS(64, 64, 1, 2)  // (the input channels split over two workgroups: the accumulators
    CFM_WGRAD_ROWS(96, 32, 1, 2)  //  of all of them do not fit in registers)
#undef CFM_WGRAD_ROWS
    return false;
}

static bool wgrad_lds(int cin, int cout, int nslab, hipStream_t st, const bf16 *X, const bf16 *dZ, int kp,
                      float *partial, int n, int npts, int slab) {
    if (wgrad_rows(cin, cout, nslab, st, X, dZ, kp, partial, n, npts, slab)) return true;
    const dim3 grid(27, nslab);
#define CFM_WGRAD_LDS(C, O)                                                                              \
    if (cin == C && cout == O) {                                                                         \
        hipLaunchKernelGGL((conv3_wgrad_lds_kernel<C, O>), grid, dim3(256), 0, st, X, dZ, kp, partial, n, \
                           npts, slab);                                                                  \
        return true;                                                                                     \
    }
    CFM_WGRAD_LDS(16, 16)
    CFM_WGRAD_LDS(16, 32)
    CFM_WGRAD_LDS(32, 32)
    CFM_WGRAD_LDS(32, 64)
    CFM_WGRAD_LDS(64, 64)
    CFM_WGRAD_LDS(96, 32)
    CFM_WGRAD_LDS(48, 16)
#undef CFM_WGRAD_LDS
    return false;
}

// out[i] = sum over the nslab slab partials of u->partial (count values each), in two passes
static void reduce_slabs(cfm_unet *u, int nslab, size_t count, float *out) {
    if (nslab < 4 * kReduceParts) {
        hipLaunchKernelGGL(reduce_slabs_kernel, blocks_for(count), dim3(256), 0, u->stream, u->partial, nslab, count,
                           out);
        return;
    }
    const dim3 grid((unsigned)((count + 255) / 256), kReduceParts);
    hipLaunchKernelGGL(reduce_slab_parts_kernel, grid, dim3(256), 0, u->stream, u->partial, nslab, count, kReduceParts,
                       u->rtmp);
    hipLaunchKernelGGL(reduce_slabs_kernel, blocks_for(count), dim3(256), 0, u->stream, u->rtmp, kReduceParts, count,
                       out);
}

static cfm_status conv_wgrad(cfm_unet *u, int I, const bf16 *X, const bf16 *dZ) {
    const ConvSpec &s = kConv[I];
    const int level = s.level, n = u->n >> level;
    const size_t npts = u->P[level];
    const size_t slab = u->slab;  // fixed 4096-point slabs measured fastest (more wavefronts)
    const int nslab = (int)((npts + slab - 1) / slab);
    const int ntile_k = u->Kp[I] / 16;
    const int ntile = (s.cout / 16) * ntile_k;
    const size_t count = (size_t)s.cout * u->Kp[I];
    if (!wgrad_lds(s.cin, s.cout, nslab, u->stream, X, dZ, u->Kp[I], u->partial, n, (int)npts, (int)slab))
        hipLaunchKernelGGL(conv3_wgrad_kernel, dim3((ntile + 3) / 4, nslab), dim3(256), 0, u->stream, X, s.cin, dZ,
                           s.cout, u->Kp[I], u->partial, n, (int)npts, (int)slab, ntile_k);
    CFM_HIP(hipGetLastError());
    reduce_slabs(u, nslab, count, u->gk);
    hipLaunchKernelGGL(unpack_wgrad_kernel, blocks_for((size_t)s.cout * s.cin_real * 27), dim3(256), 0, u->stream,
                       u->gk, s.cout, s.cin_real, s.cin, u->Kp[I], u->grad + u->woff[I]);
    hipLaunchKernelGGL(chan_sum_kernel, dim3(nslab), dim3(256), 0, u->stream, dZ, s.cout, (int)npts, (int)slab,
                       u->partial);
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(s.cout), dim3(1024), 0, u->stream, u->partial, nslab,
                       u->grad + u->boff[I]);
    CFM_HIP(hipGetLastError());
    return CFM_OK;
}

// dX = conv_transpose(dZ) for conv I (output channels = cin of the layer); with mask (the
// forward output of the ReLU layer feeding conv I), the gradient before that ReLU
static cfm_status conv_dgrad(cfm_unet *u, int I, const bf16 *dZ, bf16 *dX, const bf16 *mask = nullptr) {
    const ConvSpec &s = kConv[I];
    return conv_fwd(u, dZ, s.cout, u->Wt[I], u->Kpt[I], nullptr, dX, s.cin, s.level, 0, mask);
}

static cfm_status run_backward(cfm_unet *u, const float *target_host, double *loss) {
    const size_t P1 = u->P[0], P2 = u->P[1], P3 = u->P[2];
    const int n = u->n;
    CFM_HIP(hipMemcpyAsync(u->target, target_host, 4 * P1, hipMemcpyHostToDevice, u->stream));
    const unsigned nb1 = (unsigned)((P1 + 255) / 256);
    hipLaunchKernelGGL(mse_kernel, dim3(nb1), dim3(256), 0, u->stream, u->pred, u->target, u->dpred, u->lpart, P1);
    const float *hw = u->master + u->head_off;
    // dZ9: the head's data gradient through the last ReLU
    hipLaunchKernelGGL(head_bwd_kernel, dim3(nb1), dim3(256), 0, u->stream, u->A9, hw, u->dpred, u->g1b, u->hpart, P1);
    CFM_HIP(hipGetLastError());
    // head gradient: reduce per-block partials (fixed order)
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(17), dim3(1024), 0, u->stream, u->hpart, (int)nb1,
                       u->grad + u->head_off);
    cfm_status st;
#define CK(x) \
    if ((st = (x)) != CFM_OK) return st
    // Every ReLU backward is folded into the kernel producing the gradient it masks (the data-
    // gradient convs, the upsampling and pooling gradients, the head), and the skip gradients are
    // read in place from the concatenated buffers' gradients (g1c, g2c), which stay untouched
    // until the encoder levels use them.
    // decoder level 1
    CK(conv_wgrad(u, 9, u->A8, u->g1b));
    CK(conv_dgrad(u, 9, u->g1b, u->g1a, u->A8));       // dZ8
    CK(conv_wgrad(u, 8, u->U1, u->g1a));
    CK(conv_dgrad(u, 8, u->g1a, u->g1c));               // dU1 [P1][48]
    hipLaunchKernelGGL(up_bwd_kernel, blocks_for(P2 * 32 / 8), dim3(256), 0, u->stream, u->g1c, u->g2b, 32, 48, 0,
                       n / 2, u->batch, u->A7);          // dZ7
    // decoder level 2
    CK(conv_wgrad(u, 7, u->A6, u->g2b));
    CK(conv_dgrad(u, 7, u->g2b, u->g2a, u->A6));       // dZ6
    CK(conv_wgrad(u, 6, u->U2, u->g2a));
    CK(conv_dgrad(u, 6, u->g2a, u->g2c));               // dU2 [P2][96]
    hipLaunchKernelGGL(up_bwd_kernel, blocks_for(P3 * 64 / 8), dim3(256), 0, u->stream, u->g2c, u->g3b, 64, 96, 0,
                       n / 4, u->batch, u->A5);          // dZ5
    // bottleneck
    CK(conv_wgrad(u, 5, u->A4, u->g3b));
    CK(conv_dgrad(u, 5, u->g3b, u->g3a, u->A4));       // dZ4
    CK(conv_wgrad(u, 4, u->Q2, u->g3a));
    CK(conv_dgrad(u, 4, u->g3a, u->g3b));               // dQ2 [P3][32]
    hipLaunchKernelGGL(pool_bwd_kernel, blocks_for(P2 * 32 / 8), dim3(256), 0, u->stream, u->g3b, u->g2b, 32, n / 2,
                       u->batch, u->g2c, 96, 64, u->S2);  // dZ3 = relu'(S2) (pool' dQ2 + dS2 skip)
    // encoder level 2
    CK(conv_wgrad(u, 3, u->A2, u->g2b));
    CK(conv_dgrad(u, 3, u->g2b, u->g2a, u->A2));       // dZ2
    CK(conv_wgrad(u, 2, u->Q1, u->g2a));
    CK(conv_dgrad(u, 2, u->g2a, u->g2b));               // dQ1 [P2][16]
    hipLaunchKernelGGL(pool_bwd_kernel, blocks_for(P1 * 16 / 8), dim3(256), 0, u->stream, u->g2b, u->g1a, 16, n,
                       u->batch, u->g1c, 48, 32, u->S1);  // dZ1 = relu'(S1) (pool' dQ1 + dS1 skip)
    // encoder level 1
    CK(conv_wgrad(u, 1, u->A0, u->g1a));
    CK(conv_dgrad(u, 1, u->g1a, u->g1b, u->A0));       // dZ0
    CK(conv_wgrad(u, 0, u->X0, u->g1b));
#undef CK
    CFM_HIP(hipGetLastError());
    if (loss) {
        std::vector<double> lp(nb1);
        CFM_HIP(hipMemcpyAsync(lp.data(), u->lpart, 8 * nb1, hipMemcpyDeviceToHost, u->stream));
        CFM_HIP(hipStreamSynchronize(u->stream));
        double s = 0.0;
        for (double v : lp) s += v;
        *loss = s / (double)P1;
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_unet_gradient(cfm_unet *u, const float *input, const float *target, float *grad,
                                        double *loss) {
    if (!u || !input || !target) return CFM_EINVAL;
    cfm_status st = run_forward(u, input);
    if (st == CFM_OK) st = run_backward(u, target, loss);
    if (st != CFM_OK) return st;
    if (grad) CFM_HIP(hipMemcpyAsync(grad, u->grad, 4 * u->nparam, hipMemcpyDeviceToHost, u->stream));
    CFM_HIP(hipStreamSynchronize(u->stream));
    return CFM_OK;
}

extern "C" cfm_status cfm_unet_train_step(cfm_unet *u, const float *input, const float *target, float lr,
                                          double *loss) {
    if (!u || !input || !target || !(lr > 0.f)) return CFM_EINVAL;
    cfm_status st = run_forward(u, input);
    if (st == CFM_OK) st = run_backward(u, target, loss);
    if (st != CFM_OK) return st;
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    hipLaunchKernelGGL(adam_kernel, blocks_for(u->nparam), dim3(256), 0, u->stream, u->master, u->grad, u->adam_m,
                       u->adam_v, u->nparam, lr, b1, b2, 1e-8f, c1, c2);
    CFM_HIP(hipGetLastError());
    st = pack_all_weights(u);
    if (st != CFM_OK) return st;
    CFM_HIP(hipStreamSynchronize(u->stream));
    return CFM_OK;
}

// He-uniform initialisation from a SplitMix64 stream: w ~ U(-a, a),
// a = sqrt(6 / fan_in); biases 0; head weights scaled by 0.1.
extern "C" cfm_status cfm_unet_init_params(uint64_t seed, float *params) {
    if (!params) return CFM_EINVAL;
    uint64_t s = seed ^ 0xA0761D6478BD642FULL;
    auto uni = [&]() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    size_t off = 0;
    for (const auto &c : kConv) {
        const double a = std::sqrt(6.0 / (27.0 * c.cin_real));
        const size_t nw = (size_t)c.cout * c.cin_real * 27;
        for (size_t i = 0; i < nw; ++i) params[off + i] = (float)((2.0 * uni() - 1.0) * a);
        off += nw;
        for (int i = 0; i < c.cout; ++i) params[off + i] = 0.f;
        off += c.cout;
    }
    const double a = 0.1 * std::sqrt(6.0 / kHeadIn);
    for (int i = 0; i < kHeadIn; ++i) params[off + i] = (float)((2.0 * uni() - 1.0) * a);
    params[off + kHeadIn] = 0.f;
    return CFM_OK;
}
