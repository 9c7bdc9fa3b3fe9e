// Device pieces shared by the single-system and batched CG kernels: the
// SELL-64 matrix view, the per-row SpMV and the fixed-tree reductions.
#pragma once

#include "cfm_common.hpp"

namespace cfm {

constexpr int kChunk = 1024;
constexpr int kThreads = 256;
constexpr int kSlicesPerChunk = kChunk / 64;

// Matrix view of the CG kernels.  SELL-64 (ndiag == 0): slices of 64 rows,
// entry j of the 64 rows of slice s at slice_ptr[s] + 64 j + lane, column -1
// for padding.  DIA (ndiag > 0, banded matrices with sorted rows): diagonal d
// holds a(row, row + doff[d]) at val[d * npad + row], 0 where the CSR has no
// entry; no column indices are stored.  Both give each row's sum in the CSR's
// stored column order (a stored 0 adds an exact 0 to a non-negative-zero sum).
struct Sell {
    const int *slice_ptr;  // [nslices + 1], entry offsets (multiples of 64)
    const int *col;        // [slice_ptr[nslices]], -1 = padding
    const void *val;       // double or float
    const int *doff;       // DIA: [ndiag] ascending column offsets
    int ndiag;             // 0 = SELL-64
    int n;                 // DIA: rows (columns outside [0, n) are skipped)
    int npad;              // DIA: stride between diagonals
};


// ------------------------------------------------------------ SpMV row ----
template <typename VT, typename XT>
__device__ __forceinline__ XT sell_row(const Sell &S, const XT *__restrict__ x, int slice, int lane) {
    if (S.ndiag > 0) {  // uniform over the grid
        const int row = slice * 64 + lane;
        const VT *vp = static_cast<const VT *>(S.val) + row;
        XT acc = XT(0);
        // one diagonal at a time: an 8-wide unrolled form pushed the persistent
        // pipelined kernel into scratch memory
        for (int d = 0; d < S.ndiag; ++d) {
            const int c = row + S.doff[d];
            const VT v = vp[(size_t)d * S.npad];
            if (c >= 0 && c < S.n) acc = fma((XT)v, x[c], acc);
        }
        return acc;
    }
    const int beg = S.slice_ptr[slice];
    const int width = (S.slice_ptr[slice + 1] - beg) >> 6;
    const int *cp = S.col + beg + lane;
    const VT *vp = static_cast<const VT *>(S.val) + beg + lane;
    XT acc = XT(0);
    for (int j = 0; j < width; j += 8) {
        int c[8];
        VT v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool ok = j + u < width;
            c[u] = ok ? cp[(j + u) * 64] : -1;
            v[u] = ok ? vp[(j + u) * 64] : VT(0);
        }
        XT xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : XT(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c[u] >= 0) acc = fma((XT)v[u], xv[u], acc);
    }
    return acc;
}

// A workgroup is SUB sub-blocks of 256 threads; every sub-block forms its
// sums with the same fixed tree: xor butterfly per wavefront, then
// (w0 + w1) + (w2 + w3) over its four wavefronts.  lds holds 4*SUB doubles.
template <int SUB>
__device__ __forceinline__ double sub_sum(double v, double *lds) {
    v = wave_butterfly(v);
    __syncthreads();  // lds may still be read from a previous call
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    const int w0 = (threadIdx.x >> 8) * 4;
    return (lds[w0] + lds[w0 + 1]) + (lds[w0 + 2] + lds[w0 + 3]);
}

// Reduce nchunks partials (fixed order) -> same value in every thread: each
// sub-block's thread t adds partials t, t+256, ... and the sub-block tree
// combines them (every sub-block computes the identical value).
// The loads go out eight at a time before the adds (same order of additions, so
// the same bits): after a device barrier they come from the memory-side cache,
// and one latency per eight partials replaces one per partial.
template <int SUB>
__device__ __forceinline__ double reduce_partials(const double *part, int nchunks, double *lds) {
    double a = 0.0;
    int j = threadIdx.x & (kThreads - 1);
    for (; j + 7 * kThreads < nchunks; j += 8 * kThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[j + u * kThreads];
#pragma unroll
        for (int u = 0; u < 8; ++u) a = a + v[u];
    }
    for (; j < nchunks; j += kThreads) a = a + part[j];
    return sub_sum<SUB>(a, lds);
}

__device__ __forceinline__ bool is_finite(double v) { return v - v == 0.0; }

// Host check of a CSR before any storage builder reads it: rowptr[0] = base,
// rows non-decreasing, rowptr[n] = nnz + base, and (when col is given) every
// column in [base, ncols + base).  Records the first violation for
// cfm_last_error() and returns CFM_EINVAL; the builders then index col/val
// only inside [0, nnz).
inline cfm_status check_csr(long long n, long long nnz, const int *rowptr, const int *col, int base,
                            long long ncols, const char *what) {
    if (rowptr[0] != base)
        return CFM_FAIL(CFM_EINVAL, "%s: rowptr[0] = %d, expected the index base %d", what, rowptr[0], base);
    for (long long r = 0; r < n; ++r)
        if (rowptr[r + 1] < rowptr[r])
            return CFM_FAIL(CFM_EINVAL, "%s: rowptr decreases at row %lld (%d > %d)", what, r, rowptr[r],
                            rowptr[r + 1]);
    if ((long long)rowptr[n] - base != nnz)
        return CFM_FAIL(CFM_EINVAL, "%s: rowptr[n] - base = %lld, nnz = %lld", what, (long long)rowptr[n] - base,
                        nnz);
    if (col)
        for (long long j = 0; j < nnz; ++j)
            if ((long long)col[j] < base || (long long)col[j] >= ncols + base)
                return CFM_FAIL(CFM_EINVAL, "%s: column %d of entry %lld is outside [%d, %lld)", what, col[j], j,
                                base, ncols + base);
    return CFM_OK;
}

}  // namespace cfm
