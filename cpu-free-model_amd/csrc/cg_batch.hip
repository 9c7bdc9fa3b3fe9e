// Batched CG: many independent SPD systems (a block-diagonal CSR) solved in
// ONE persistent cooperative launch, each system to its own tolerance.
//
// Method: Pipelined CG (one device-wide barrier per iteration), the
// recurrences of CG/src/common.cu:609-780 / CG/src/single-stream/
// pipelined.cu:240-317 applied per system with per-system scalars.
//
// Layout: every system's rows start on a 1024-row chunk boundary, so each
// system's dot products use exactly the chunk tree of a single-system solve:
// a batch member reproduces the solo solve of the same system bit for bit.
//
// Scheduling: at every iteration the chunks of the still-active systems are
// numbered in system order and cut into one contiguous range per workgroup
// (load balance follows convergence: a converged system frees its share).
// A workgroup reduces the dot products of the systems its range touches; the
// workgroup whose range holds a system's first chunk ("owner") writes that
// system's scalars, history and status.  Scalars carried to the next
// iteration (alpha, delta) are double-buffered by iteration parity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cg_common.hpp"
#include "unet.hpp"

namespace cfm {

constexpr int kMaxSys = 4096;
constexpr int kMaxOverlap = 64;  // systems handled per pass of a workgroup's range

enum { kActive = 0, kConverged = 1, kBreakdown = 2, kMaxIter = 3 };

struct BatchArgs {
    int nsys;
    int total_chunks;
    const int *sys_chunk0;  // [nsys+1] first chunk of each system
    const int *sys_rows;    // [nsys] real rows of each system
    Sell A;                 // SELL-64 over the padded global rows
    const double *b;
    const double *x0;
    double *x, *r, *z, *s, *p;
    double *w[2];
    double *part;      // [5][total_chunks]: delta buf0/1, gamma buf0/1, <b,b>
    double *st_alpha;  // [2][nsys]
    double *st_delta;  // [2][nsys]
    int *fin;          // [nsys] iteration at which the system finished (INT_MAX: running)
    int *status;       // [nsys] final state (kConverged, kBreakdown, kMaxIter)
    int *iters;        // [nsys]
    double *hist;      // [nsys][hist_cap] or null
    int hist_cap;
    int max_iter;
    double tol;
    int quirk;
    int x0_guard;      // restart a system from x0 = 0 when <r0,r0> > <b,b>
    int *x0_rejected;  // [nsys] 1 where the guard restarted the system
    // DIA storage (dmax > 0): a(row, row + doff[sy*dmax + d]) at dval[d*dpad + row]
    const double *dval;
    const int *doff;
    int dmax;
    long long dpad;
    GridSync *sync;
};

constexpr int kNoDiag = 1 << 30;  // offset of an unused diagonal slot: never inside a system

// Row sum of A x for padded row `row` of system sy (rows [row0, row0+nrows)).
__device__ __forceinline__ double batch_row(const BatchArgs &A, const double *__restrict__ x, int slice, int lane,
                                            int row, int row0, int nrows, int sy) {
    if (A.dmax > 0) {  // uniform over the grid
        const int *off = A.doff + (size_t)sy * A.dmax;
        const double *vp = A.dval + row;
        double acc = 0.0;
        for (int d = 0; d < A.dmax; ++d) {
            const int c = row + off[d];
            const double v = vp[(size_t)d * A.dpad];
            if (c >= row0 && c - row0 < nrows) acc = fma(v, x[c], acc);
        }
        return acc;
    }
    return sell_row<double, double>(A.A, x, slice, lane);
}

struct BatchShared {
    int act_prefix[kMaxSys + 1];  // chunk offsets of active systems, in system order
    int act_ids[kMaxSys];         // active system ids in order
    int n_act;
    int ov_sys[kMaxOverlap];      // systems of the current pass
    int ov_pos0;                  // position (in act_ids) of ov_sys[0]
    double ov_alpha[kMaxOverlap];
    double ov_beta[kMaxOverlap];
    int ov_live[kMaxOverlap];
    double red[16];               // sub_sum scratch (4 * SUB)
};

// Per-row pipelined update of chunk c (system-local rows < nrows).
template <int SUB>
__device__ __forceinline__ void batch_chunk(const BatchArgs &A, int c, int sy, int row0_sys, int nrows, int buf,
                                            double alpha, double beta, bool live, double *red) {
    const int t = threadIdx.x & (kThreads - 1);
    const double *wc = A.w[buf];
    double *wn = A.w[buf ^ 1];
    const int nb = buf ^ 1;
    double ad = 0.0, ag = 0.0;
    if (live) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int slice = c * kSlicesPerChunk + i * 4 + (t >> 6);
            const int row = c * kChunk + i * kThreads + t;
            if (row - row0_sys < nrows) {
                const double qi = batch_row(A, wc, slice, t & 63, row, row0_sys, nrows, sy);
                const double wi = wc[row];
                const double zi = fma(beta, A.z[row], qi);
                const double si = fma(beta, A.s[row], wi);
                const double pi = fma(beta, A.p[row], A.r[row]);
                A.x[row] = fma(alpha, pi, A.x[row]);
                const double ri = fma(-alpha, si, A.r[row]);
                const double wn_i = fma(-alpha, zi, wi);
                A.z[row] = zi;
                A.s[row] = si;
                A.p[row] = pi;
                A.r[row] = ri;
                wn[row] = wn_i;
                ad = fma(ri, ri, ad);
                ag = fma(ri, wn_i, ag);
            }
        }
    }
    const double sd = sub_sum<SUB>(ad, red);
    const double sg = sub_sum<SUB>(ag, red);
    if (live && t == 0) {
        A.part[(0 + nb) * A.total_chunks + c] = sd;
        A.part[(2 + nb) * A.total_chunks + c] = sg;
    }
}

// system owning global chunk c (binary search over sys_chunk0)
__device__ __forceinline__ int system_of_chunk(const BatchArgs &A, int c) {
    int lo = 0, hi = A.nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.sys_chunk0[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int SUB>
__global__ void __launch_bounds__(256 * SUB) cg_batch_kernel(BatchArgs A) {
    __shared__ BatchShared S;
    const int t = threadIdx.x & (kThreads - 1);
    const int sub = threadIdx.x >> 8;
    unsigned int epoch = 0;

    // ---- init 1: r = b - A x0, x = x0, z = s = p = 0, partials <b,b>, <r0,r0> (all chunks)
    for (int base = blockIdx.x * SUB; base < A.total_chunks; base += gridDim.x * SUB) {
        const int c = base + sub;
        const bool act = c < A.total_chunks;
        int sy = 0, row0 = 0, nrows = 0;
        if (act) {
            sy = system_of_chunk(A, c);
            row0 = A.sys_chunk0[sy] * kChunk;
            nrows = A.sys_rows[sy];
        }
        double ab = 0.0, ar = 0.0;
        if (act) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int slice = c * kSlicesPerChunk + i * 4 + (t >> 6);
                const int row = c * kChunk + i * kThreads + t;
                if (row - row0 < nrows) {
                    const double ax = batch_row(A, A.x0, slice, t & 63, row, row0, nrows, sy);
                    const double bi = A.b[row];
                    const double ri = fma(-1.0, ax, bi);
                    A.r[row] = ri;
                    A.x[row] = A.x0[row];
                    A.z[row] = 0.0;
                    A.s[row] = 0.0;
                    A.p[row] = 0.0;
                    ab = fma(bi, bi, ab);
                    ar = fma(ri, ri, ar);
                }
            }
        }
        const double sb = sub_sum<SUB>(ab, S.red);
        const double sr = sub_sum<SUB>(ar, S.red);
        if (act && t == 0) {
            A.part[4 * A.total_chunks + c] = sb;
            A.part[1 * A.total_chunks + c] = sr;  // delta buffer 1 is free until iteration 0 writes it
        }
    }
    if (!grid_barrier(A.sync, epoch)) return;
    // ---- init 2: guard decision, w0 = A r, partials <r,r>, <r,w0> into buffer 0.
    // With the guard on, every chunk of a system reduces that system's <r0,r0>
    // and <b,b> with the same tree, so all of them reach the same decision; a
    // rejected system takes x = 0, r = b and w0 = A b (b is read-only, so no
    // chunk reads a residual another chunk is rewriting).
    for (int base = blockIdx.x * SUB; base < A.total_chunks; base += gridDim.x * SUB) {
        const int c = base + sub;
        const bool act = c < A.total_chunks;
        int sy = 0, row0 = 0, nrows = 0;
        if (act) {
            sy = system_of_chunk(A, c);
            row0 = A.sys_chunk0[sy] * kChunk;
            nrows = A.sys_rows[sy];
        }
        bool reject = false;
        if (A.x0_guard) {  // uniform over the workgroup
            double a1 = 0.0, a3 = 0.0;
            if (act) {
                const int c0 = A.sys_chunk0[sy], nc = A.sys_chunk0[sy + 1] - c0;
                for (int j = t; j < nc; j += kThreads) {
                    a1 = a1 + A.part[1 * A.total_chunks + c0 + j];
                    a3 = a3 + A.part[4 * A.total_chunks + c0 + j];
                }
            }
            const double d0 = sub_sum<SUB>(a1, S.red);
            const double bb = sub_sum<SUB>(a3, S.red);
            reject = act && !(d0 <= bb);  // a non-finite residual is rejected too
            if (act && t == 0 && c == A.sys_chunk0[sy]) A.x0_rejected[sy] = reject ? 1 : 0;
        }
        double ad = 0.0, ag = 0.0;
        if (act) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int slice = c * kSlicesPerChunk + i * 4 + (t >> 6);
                const int row = c * kChunk + i * kThreads + t;
                if (row - row0 < nrows) {
                    double wi, ri;
                    if (reject) {
                        wi = batch_row(A, A.b, slice, t & 63, row, row0, nrows, sy);
                        ri = A.b[row];
                        A.r[row] = ri;
                        A.x[row] = 0.0;
                    } else {
                        wi = batch_row(A, A.r, slice, t & 63, row, row0, nrows, sy);
                        ri = A.r[row];
                    }
                    A.w[0][row] = wi;
                    ad = fma(ri, ri, ad);
                    ag = fma(ri, wi, ag);
                }
            }
        }
        const double sd = sub_sum<SUB>(ad, S.red);
        const double sg = sub_sum<SUB>(ag, S.red);
        if (act && t == 0) {
            A.part[0 * A.total_chunks + c] = sd;
            A.part[2 * A.total_chunks + c] = sg;
        }
    }
    if (!grid_barrier(A.sync, epoch)) return;

    for (int k = 0;; ++k) {
        const int buf = k & 1;
        // ---- active list: systems with fin >= k (fin[s] = iteration at which s
        // finished, written once by its owner; a read before or after the owner's
        // write in this same iteration gives the same answer), in system order.
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            int run = 0, nact = 0;
            for (int b0 = 0; b0 < A.nsys; b0 += 64) {
                const int sy = b0 + lane;
                const int cnt = (sy < A.nsys && A.fin[sy] >= k) ? A.sys_chunk0[sy + 1] - A.sys_chunk0[sy] : 0;
                const int flag = cnt > 0 ? 1 : 0;
                int incl = cnt, fincl = flag;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const int v = __shfl_up(incl, off, 64);
                    const int f = __shfl_up(fincl, off, 64);
                    if (lane >= off) {
                        incl += v;
                        fincl += f;
                    }
                }
                if (flag) {
                    const int pos = nact + fincl - 1;
                    S.act_ids[pos] = sy;
                    S.act_prefix[pos] = run + incl - cnt;
                }
                run += __shfl(incl, 63, 64);
                nact += __shfl(fincl, 63, 64);
            }
            if (lane == 0) {
                S.n_act = nact;
                S.act_prefix[nact] = run;
            }
        }
        __syncthreads();
        const int n_act = S.n_act;
        const int total = S.act_prefix[n_act];
        if (total == 0) break;  // every system finished: uniform over the grid
        const int per = (total + gridDim.x - 1) / gridDim.x;
        const int lo = min(total, (int)blockIdx.x * per);
        const int hi = min(total, lo + per);

        // position (in act_ids) of the system holding active chunk index a
        auto pos_of = [&](int a) {
            int l = 0, h = n_act - 1;
            while (l < h) {
                const int mid = (l + h + 1) >> 1;
                if (S.act_prefix[mid] <= a) l = mid;
                else h = mid - 1;
            }
            return l;
        };
        int pos = lo < hi ? pos_of(lo) : n_act;
        const int pos_end = lo < hi ? pos_of(hi - 1) + 1 : n_act;
        while (pos < pos_end) {  // passes of up to kMaxOverlap systems (uniform)
            const int npass = min(kMaxOverlap, pos_end - pos);
            // scalars of the pass's systems: sub-block j reduces systems j, j+SUB, ...
            for (int q0 = 0; q0 < npass; q0 += SUB) {
                const int q = q0 + sub;
                const bool has = q < npass;
                const int sy = has ? S.act_ids[pos + q] : 0;
                const int c0 = has ? A.sys_chunk0[sy] : 0;
                const int nc = has ? A.sys_chunk0[sy + 1] - c0 : 0;
                double a1 = 0.0, a2 = 0.0, a3 = 0.0;
                for (int j = t; j < nc; j += kThreads) {
                    a1 = a1 + A.part[(0 + buf) * A.total_chunks + c0 + j];
                    a2 = a2 + A.part[(2 + buf) * A.total_chunks + c0 + j];
                    a3 = a3 + A.part[4 * A.total_chunks + c0 + j];
                }
                const double delta = sub_sum<SUB>(a1, S.red);
                const double gam = sub_sum<SUB>(a2, S.red);
                const double bb = sub_sum<SUB>(a3, S.red);
                if (has && t == 0) {
                    const int first = S.act_prefix[pos + q];
                    const bool owner = first >= lo && first < hi;
                    const double thr = A.tol > 0.0 ? A.tol * A.tol * bb : 0.0;
                    int st = kActive;
                    double alpha = 0.0, beta = 0.0;
                    if (owner && A.hist && k < A.hist_cap) A.hist[(size_t)sy * A.hist_cap + k] = delta;
                    if (delta <= thr) {
                        st = kConverged;
                    } else if (k >= A.max_iter) {
                        st = kMaxIter;
                    } else {
                        double denom;
                        if (A.quirk ? (k > 1) : (k > 0)) {
                            beta = delta / A.st_delta[(buf ^ 1) * A.nsys + sy];
                            denom = gam - (beta / A.st_alpha[(buf ^ 1) * A.nsys + sy]) * delta;
                        } else {
                            beta = 0.0;
                            denom = gam;
                        }
                        if (!(denom > 0.0) || !is_finite(denom)) {
                            st = kBreakdown;
                        } else {
                            alpha = delta / denom;
                        }
                    }
                    if (owner) {
                        if (st == kActive) {
                            A.st_alpha[buf * A.nsys + sy] = alpha;
                            A.st_delta[buf * A.nsys + sy] = delta;
                        } else {
                            A.fin[sy] = k;
                            A.status[sy] = st;
                            A.iters[sy] = k;
                        }
                    }
                    S.ov_sys[q] = sy;
                    S.ov_alpha[q] = alpha;
                    S.ov_beta[q] = beta;
                    S.ov_live[q] = st == kActive;
                }
            }
            __syncthreads();
            // chunks of this pass inside [lo, hi)
            const int a_lo = max(lo, S.act_prefix[pos]);
            const int a_hi = min(hi, S.act_prefix[pos + npass]);
            for (int a0 = a_lo; a0 < a_hi; a0 += SUB) {
                const int a = a0 + sub;
                const bool act = a < a_hi;
                int q = 0;
                if (act) {
                    q = 0;
                    while (q + 1 < npass && S.act_prefix[pos + q + 1] <= a) ++q;
                }
                const int sy = S.ov_sys[q];
                const int c = A.sys_chunk0[sy] + (a - S.act_prefix[pos + q]);
                batch_chunk<SUB>(A, c, sy, A.sys_chunk0[sy] * kChunk, A.sys_rows[sy], buf, S.ov_alpha[q],
                                 S.ov_beta[q], act && S.ov_live[q], S.red);
            }
            __syncthreads();
            pos += npass;
        }
        if (!grid_barrier(A.sync, epoch)) return;
    }
}

// x0[chunk-aligned row of system first + j / n3] = scale * pred[j], j < count * n3
__global__ void scatter_x0_kernel(const float *__restrict__ pred, const int *__restrict__ sys_chunk0, int first,
                                  int count, int n3, double scale, double *__restrict__ x0) {
    const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j >= (size_t)count * n3) return;
    const int q = (int)(j / n3), i = (int)(j % n3);
    x0[(size_t)sys_chunk0[first + q] * kChunk + i] = scale * (double)pred[j];
}

// One damped Jacobi sweep: xo = x + omega * (b - A x) / a_ii.  (A x)_i is
// the SpMV row sum (fma in stored column order), then r = fma(-1, Ax, b),
// xo = fma(omega, r / a_ii, x).  Rows without a stored diagonal (the chunk
// padding) keep x.
__global__ void jacobi_smooth_kernel(Sell A, const double *__restrict__ b, const double *__restrict__ x,
                                     double *__restrict__ xo, int nrows, double omega) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const int slice = row >> 6, lane = row & 63;
    const int beg = A.slice_ptr[slice];
    const int width = (A.slice_ptr[slice + 1] - beg) >> 6;
    const double *val = static_cast<const double *>(A.val);
    double ax = 0.0, d = 0.0;
    for (int j = 0; j < width; ++j) {
        const int c = A.col[beg + j * 64 + lane];
        if (c >= 0) {
            const double v = val[beg + j * 64 + lane];
            ax = fma(v, x[c], ax);
            if (c == row) d = v;
        }
    }
    if (d != 0.0) {
        const double r = fma(-1.0, ax, b[row]);
        xo[row] = fma(omega, r / d, x[row]);
    } else {
        xo[row] = x[row];
    }
}

// The same sweep on DIA storage: row -> system by binary search over the
// chunk table; padding rows keep x.
__global__ void jacobi_smooth_dia_kernel(const double *__restrict__ dval, const int *__restrict__ doff, int dmax,
                                         long long dpad, const int *__restrict__ sys_chunk0,
                                         const int *__restrict__ sys_rows, int nsys, const double *__restrict__ b,
                                         const double *__restrict__ x, double *__restrict__ xo, int nrows,
                                         double omega) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const int c = row / kChunk;
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys_chunk0[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    const int row0 = sys_chunk0[lo] * kChunk, nr = sys_rows[lo];
    if (row - row0 >= nr) {
        xo[row] = x[row];
        return;
    }
    const int *off = doff + (size_t)lo * dmax;
    double ax = 0.0, d = 0.0;
    for (int k = 0; k < dmax; ++k) {
        const int cc = row + off[k];
        const double v = dval[(size_t)k * dpad + row];
        if (cc >= row0 && cc - row0 < nr) {
            ax = fma(v, x[cc], ax);
            if (off[k] == 0) d = v;
        }
    }
    if (d != 0.0) {
        const double r = fma(-1.0, ax, b[row]);
        xo[row] = fma(omega, r / d, x[row]);
    } else {
        xo[row] = x[row];
    }
}

}  // namespace cfm

using namespace cfm;

struct cfm_cg_batch_plan {
    cfm_cg_params p{};
    DeviceProps dev{};
    int nsys = 0, total_chunks = 0;
    long long total_rows = 0;
    std::vector<int> h_chunk0, h_rows;
    std::vector<long long> h_row0;  // caller's global row offsets
    int *sys_chunk0 = nullptr, *sys_rows = nullptr;
    int *slice_ptr = nullptr, *scol = nullptr;
    double *sval = nullptr;
    double *b = nullptr, *x0 = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *s = nullptr, *pv = nullptr;
    double *w[2] = {nullptr, nullptr};
    double *part = nullptr, *st_alpha = nullptr, *st_delta = nullptr, *hist = nullptr;
    int *fin = nullptr, *status = nullptr, *iters = nullptr, *x0_rejected = nullptr;
    double *dval = nullptr;  // DIA storage (dmax > 0 when built)
    int *doff = nullptr;
    int dmax = 0;
    bool dia_on = false;  // solve on DIA (else SELL-64); both are built for an auto plan of banded systems
    float *features = nullptr;  // warm-start network input [nsys][2][rows], loaded on demand
    int hist_cap = 0;
    GridSync *sync = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int sub = 4, grid = 0;
    unsigned long long spin_ticks = 0;
};

using BatchKernelFn = void (*)(BatchArgs);
static BatchKernelFn batch_kernel(int sub) {
    switch (sub) {
        case 1: return cg_batch_kernel<1>;
        case 2: return cg_batch_kernel<2>;
        case 4: return cg_batch_kernel<4>;
        default: return nullptr;
    }
}

extern "C" void cfm_cg_batch_plan_destroy(cfm_cg_batch_plan *pl) {
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    void *ptrs[] = {pl->sys_chunk0, pl->sys_rows, pl->slice_ptr, pl->scol, pl->sval, pl->b, pl->x0, pl->x,
                    pl->r, pl->z, pl->s, pl->pv, pl->w[0], pl->w[1], pl->part, pl->st_alpha, pl->st_delta,
                    pl->hist, pl->fin, pl->status, pl->iters, pl->x0_rejected, pl->features, pl->dval, pl->doff,
                    pl->sync};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (pl->ev0) (void)hipEventDestroy(pl->ev0);
    if (pl->ev1) (void)hipEventDestroy(pl->ev1);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

static cfm_status batch_configure(cfm_cg_batch_plan *pl, int sub, int grid) {
    BatchKernelFn k = batch_kernel(sub);
    if (!k) return CFM_EINVAL;
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, 256 * sub, 0);
    if (e != hipSuccess || per_cu < 1)
        return CFM_FAIL(CFM_EDEVICE, "batched CG: occupancy query for %d threads gave %d workgroups per CU (%s)",
                        256 * sub, per_cu, hipGetErrorString(e));
    const int max_grid = std::min(per_cu * pl->dev.cus, kMaxGrid);
    pl->sub = sub;
    pl->grid = std::max(1, std::min(grid > 0 ? grid : max_grid, max_grid));
    return CFM_OK;
}

// DIA storage for the whole batch when every system is banded: at most
// kMaxDiagBatch distinct column offsets, rows sorted by column, zero fill at
// most half the system's nnz.  offs[sy*dmax + d] ascending, unused slots
// kNoDiag; dv[d*padded_rows + padded row].  Returns false to use SELL-64.
constexpr int kMaxDiagBatch = 16;
static bool batch_dia(int nsys, const int64_t *srp, const int *rowptr, const int *col, const double *val,
                      const std::vector<int> &chunk0, long long padded_rows, int &dmax, std::vector<int> &offs,
                      std::vector<double> &dv) {
    std::vector<std::vector<int>> so(nsys);
    dmax = 0;
    for (int sy = 0; sy < nsys; ++sy) {
        const long long g0 = srp[sy], g1 = srp[sy + 1];
        auto &o = so[sy];
        for (long long r = g0; r < g1; ++r) {
            long long prev = -1;
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const long long c = col[j];
                if (c < g0 || c >= g1 || c <= prev) return false;  // outside the block, unsorted or repeated
                prev = c;
                const int d = (int)(c - r);
                if (std::find(o.begin(), o.end(), d) == o.end()) {
                    if ((int)o.size() == kMaxDiagBatch) return false;
                    o.push_back(d);
                }
            }
        }
        const long long nnz = rowptr[g1] - rowptr[g0];
        if ((long long)o.size() * (g1 - g0) > nnz + nnz / 2) return false;
        std::sort(o.begin(), o.end());
        dmax = std::max(dmax, (int)o.size());
    }
    if (dmax == 0) return false;
    offs.assign((size_t)nsys * dmax, kNoDiag);
    for (int sy = 0; sy < nsys; ++sy)
        for (size_t d = 0; d < so[sy].size(); ++d) offs[(size_t)sy * dmax + d] = so[sy][d];
    dv.assign((size_t)dmax * padded_rows, 0.0);
    for (int sy = 0; sy < nsys; ++sy) {
        const long long g0 = srp[sy], g1 = srp[sy + 1];
        const long long prow0 = (long long)chunk0[sy] * kChunk;
        const auto &o = so[sy];
        for (long long r = g0; r < g1; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const size_t d = std::lower_bound(o.begin(), o.end(), (int)(col[j] - r)) - o.begin();
                dv[d * padded_rows + prow0 + (r - g0)] = val[j];
            }
    }
    return true;
}

extern "C" cfm_status cfm_cg_batch_plan_create(int nsys, const int64_t *sys_row_ptr, const int *rowptr,
                                               const int *col, const double *val, const cfm_cg_params *p,
                                               cfm_cg_batch_plan **out) {
    if (!out) return CFM_FAIL(CFM_EINVAL, "batched CG: null output handle");
    *out = nullptr;
    if (nsys < 1 || nsys > kMaxSys || !sys_row_ptr || !rowptr || !col || !val || !p)
        return CFM_FAIL(CFM_EINVAL, "batched CG: need 1..%d systems and non-null arrays (nsys %d)", kMaxSys, nsys);
    if (p->max_iter < 0 || p->method != CFM_CG_PIPELINED || p->mode != CFM_CG_PERSISTENT)
        return CFM_FAIL(CFM_EINVAL, "batched CG: Pipelined persistent method and max_iter >= 0 required");
    if (sys_row_ptr[0] != 0) return CFM_FAIL(CFM_EINVAL, "batched CG: sys_row_ptr[0] must be 0");
    for (int sy = 0; sy < nsys; ++sy)
        if (sys_row_ptr[sy + 1] <= sys_row_ptr[sy] || sys_row_ptr[sy + 1] - sys_row_ptr[sy] > 0x3fffffff)
            return CFM_FAIL(CFM_EINVAL, "batched CG: system %d has %lld rows", sy,
                            (long long)(sys_row_ptr[sy + 1] - sys_row_ptr[sy]));
    // rows non-decreasing before any builder reads col/val (columns are checked per system below)
    if (sys_row_ptr[nsys] > 0x7fffffffLL)
        return CFM_FAIL(CFM_EINVAL, "batched CG: %lld rows exceed 2^31", (long long)sys_row_ptr[nsys]);
    {
        const long long rows = sys_row_ptr[nsys];
        const cfm_status vst = check_csr(rows, (long long)rowptr[rows] - rowptr[0], rowptr, nullptr, 0, rows,
                                         "batched CG");
        if (vst != CFM_OK) return vst;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CFM_FAIL(CFM_EDEVICE, "no GPU");
    auto *pl = new (std::nothrow) cfm_cg_batch_plan();
    if (!pl) return CFM_ENOMEM;
    auto fail = [&](cfm_status s) {
        cfm_cg_batch_plan_destroy(pl);
        return s;
    };
    pl->p = *p;
    pl->nsys = nsys;
    if (current_device_props(&pl->dev) != CFM_OK) return fail(CFM_EDEVICE);
    // chunk-aligned layout
    pl->h_chunk0.assign(nsys + 1, 0);
    pl->h_rows.assign(nsys, 0);
    pl->h_row0.assign(sys_row_ptr, sys_row_ptr + nsys + 1);
    long long chunks = 0;
    for (int sy = 0; sy < nsys; ++sy) {
        const long long nr = sys_row_ptr[sy + 1] - sys_row_ptr[sy];
        if (nr < 1 || nr > 0x3fffffff) return fail(CFM_FAIL(CFM_EINVAL, "batched CG: system %d has %lld rows", sy, nr));
        pl->h_rows[sy] = (int)nr;
        pl->h_chunk0[sy] = (int)chunks;
        chunks += (nr + kChunk - 1) / kChunk;
        if (chunks * kChunk > 0x7fffffffLL) return fail(CFM_FAIL(CFM_EINVAL, "batched CG: padded rows exceed 2^31"));
    }
    pl->h_chunk0[nsys] = (int)chunks;
    pl->total_chunks = (int)chunks;
    pl->total_rows = sys_row_ptr[nsys];
    const long long padded_rows = chunks * kChunk;
    if (p->matrix_format < 0 || p->matrix_format > 2) return fail(CFM_FAIL(CFM_EINVAL, "batched CG: matrix_format"));
    // DIA when asked (every system must be banded), and beside SELL-64 for an auto
    // plan of banded systems, where the autotuner picks one; solves start on SELL-64
    // (for the 256-system 32-96 workload DIA measured 1301 vs 1133 ms, its per-diagonal
    // loop exposing load latency that the 8-wide SELL loop hides)
    std::vector<int> doffs;
    std::vector<double> dvals;
    bool dia = false;
    if (p->matrix_format != 1) {
        dia = batch_dia(nsys, sys_row_ptr, rowptr, col, val, pl->h_chunk0, padded_rows, pl->dmax, doffs, dvals);
        if (!dia && p->matrix_format == 2)
            return fail(CFM_FAIL(CFM_ENOTSUP, "batched CG: DIA storage needs every system banded (<= %d diagonals, "
                                              "sorted rows)", kMaxDiagBatch));
        if (!dia) pl->dmax = 0;
    }
    pl->dia_on = p->matrix_format == 2;
    const bool sell = p->matrix_format != 2;
    // SELL-64 over padded rows; columns remapped, must stay inside the system
    const long long nslices = sell ? padded_rows / 64 : 0;
    std::vector<int> sp(nslices + 1, 0);
    long long tot = 0;
    for (int sy = 0; sy < nsys && sell; ++sy) {
        const long long g0 = sys_row_ptr[sy];
        const int nr = pl->h_rows[sy];
        const long long s0 = (long long)pl->h_chunk0[sy] * kSlicesPerChunk;
        const long long s1 = (long long)pl->h_chunk0[sy + 1] * kSlicesPerChunk;
        for (long long sl = s0; sl < s1; ++sl) {
            int wdt = 0;
            for (int l = 0; l < 64; ++l) {
                const long long lr = (sl - s0) * 64 + l;
                if (lr < nr) wdt = std::max(wdt, rowptr[g0 + lr + 1] - rowptr[g0 + lr]);
            }
            sp[sl] = (int)tot;
            tot += 64LL * wdt;
            if (tot > 0x7fffffffLL) return fail(CFM_FAIL(CFM_EINVAL, "batched CG: SELL storage exceeds 2^31 entries"));
        }
    }
    sp[nslices] = (int)tot;
    std::vector<int> sc((size_t)tot, -1);
    std::vector<double> sv((size_t)tot, 0.0);
    for (int sy = 0; sy < nsys && sell; ++sy) {
        const long long g0 = sys_row_ptr[sy];
        const long long g1 = sys_row_ptr[sy + 1];
        const long long prow0 = (long long)pl->h_chunk0[sy] * kChunk;
        for (long long lr = 0; lr < g1 - g0; ++lr) {
            const long long prow = prow0 + lr;
            const long long sl = prow / 64;
            const int lane = (int)(prow % 64);
            for (int j = rowptr[g0 + lr]; j < rowptr[g0 + lr + 1]; ++j) {
                const long long cg = col[j];
                if (cg < g0 || cg >= g1)  // not block diagonal
                    return fail(CFM_FAIL(CFM_EINVAL, "batched CG: row %lld of system %d has column %lld outside it",
                                         g0 + lr, sy, cg));
                const size_t e = (size_t)sp[sl] + (size_t)(j - rowptr[g0 + lr]) * 64 + lane;
                sc[e] = (int)(prow0 + (cg - g0));
                sv[e] = val[j];
            }
        }
    }
    auto dalloc = [&](void **ptr, size_t bytes) { return hipMalloc(ptr, bytes > 0 ? bytes : 16) == hipSuccess; };
    const size_t nv = (size_t)padded_rows;
    bool ok = dalloc((void **)&pl->sys_chunk0, sizeof(int) * (nsys + 1)) &&
              dalloc((void **)&pl->sys_rows, sizeof(int) * nsys) &&
              dalloc((void **)&pl->slice_ptr, sizeof(int) * sp.size()) &&
              dalloc((void **)&pl->scol, sizeof(int) * sc.size()) &&
              dalloc((void **)&pl->sval, sizeof(double) * sv.size()) &&
              dalloc((void **)&pl->part, sizeof(double) * 5 * chunks) &&
              dalloc((void **)&pl->st_alpha, sizeof(double) * 2 * nsys) &&
              dalloc((void **)&pl->st_delta, sizeof(double) * 2 * nsys) &&
              dalloc((void **)&pl->status, sizeof(int) * nsys) && dalloc((void **)&pl->iters, sizeof(int) * nsys) &&
              dalloc((void **)&pl->fin, sizeof(int) * nsys) &&
              dalloc((void **)&pl->x0_rejected, sizeof(int) * nsys) &&
              dalloc((void **)&pl->dval, sizeof(double) * dvals.size()) &&
              dalloc((void **)&pl->doff, sizeof(int) * doffs.size()) &&
              dalloc((void **)&pl->sync, sizeof(GridSync));
    double **vecs[] = {&pl->b, &pl->x0, &pl->x, &pl->r, &pl->z, &pl->s, &pl->pv, &pl->w[0], &pl->w[1]};
    for (double **v : vecs) ok = ok && dalloc((void **)v, sizeof(double) * nv);
    if (!ok) return fail(CFM_FAIL(CFM_ENOMEM, "batched CG: device allocation failed"));
    if (hipMemcpy(pl->sys_chunk0, pl->h_chunk0.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->sys_rows, pl->h_rows.data(), sizeof(int) * nsys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->slice_ptr, sp.data(), sizeof(int) * sp.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->scol, sc.data(), sizeof(int) * sc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (dia && hipMemcpy(pl->dval, dvals.data(), sizeof(double) * dvals.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (dia && hipMemcpy(pl->doff, doffs.data(), sizeof(int) * doffs.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(pl->sval, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(CFM_EDEVICE);
    for (double **v : vecs)
        if (hipMemset(*v, 0, sizeof(double) * nv) != hipSuccess) return fail(CFM_EDEVICE);
    if (hipStreamCreate(&pl->stream) != hipSuccess || hipEventCreate(&pl->ev0) != hipSuccess ||
        hipEventCreate(&pl->ev1) != hipSuccess)
        return fail(CFM_EDEVICE);
    const int bt = p->block_threads > 0 ? p->block_threads : 1024;
    if (bt % 256 || bt > 1024) return fail(CFM_FAIL(CFM_EINVAL, "batched CG: block_threads %d", bt));
    const cfm_status cs = batch_configure(pl, bt / 256, p->grid_blocks);
    if (cs != CFM_OK) return fail(cs);
    int rate_khz = 0;
    const unsigned long long us = p->spin_limit_us > 0 ? p->spin_limit_us : 10000000ull;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, pl->dev.device) != hipSuccess ||
        rate_khz <= 0)
        rate_khz = 100000;
    pl->spin_ticks = us * (unsigned long long)rate_khz / 1000ull;
    *out = pl;
    return CFM_OK;
}

// copy between the caller's packed layout and the chunk-aligned device layout
static cfm_status batch_scatter(cfm_cg_batch_plan *pl, double *dev, const double *host) {
    for (int sy = 0; sy < pl->nsys; ++sy) {
        const size_t off = (size_t)pl->h_chunk0[sy] * kChunk;
        if (host)
            CFM_HIP(hipMemcpy(dev + off, host + pl->h_row0[sy], sizeof(double) * pl->h_rows[sy], hipMemcpyHostToDevice));
        else
            CFM_HIP(hipMemset(dev + off, 0, sizeof(double) * pl->h_rows[sy]));
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_batch_set_rhs(cfm_cg_batch_plan *pl, const double *b) {
    if (!pl || !b) return CFM_EINVAL;
    return batch_scatter(pl, pl->b, b);
}

extern "C" cfm_status cfm_cg_batch_set_x0(cfm_cg_batch_plan *pl, const double *x0) {
    if (!pl) return CFM_EINVAL;
    return batch_scatter(pl, pl->x0, x0);
}

extern "C" cfm_status cfm_cg_batch_get_x(cfm_cg_batch_plan *pl, double *x) {
    if (!pl || !x) return CFM_EINVAL;
    for (int sy = 0; sy < pl->nsys; ++sy) {
        const size_t off = (size_t)pl->h_chunk0[sy] * kChunk;
        CFM_HIP(hipMemcpy(x + pl->h_row0[sy], pl->x + off, sizeof(double) * pl->h_rows[sy], hipMemcpyDeviceToHost));
    }
    return CFM_OK;
}

static cfm_status batch_run(cfm_cg_batch_plan *pl, int max_iter, double tol, int hist_cap, float *ms,
                            int *coop) {
    BatchArgs A{};
    A.nsys = pl->nsys;
    A.total_chunks = pl->total_chunks;
    A.sys_chunk0 = pl->sys_chunk0;
    A.sys_rows = pl->sys_rows;
    A.A = Sell{pl->slice_ptr, pl->scol, pl->sval};
    A.b = pl->b;
    A.x0 = pl->x0;
    A.x = pl->x;
    A.r = pl->r;
    A.z = pl->z;
    A.s = pl->s;
    A.p = pl->pv;
    A.w[0] = pl->w[0];
    A.w[1] = pl->w[1];
    A.part = pl->part;
    A.st_alpha = pl->st_alpha;
    A.st_delta = pl->st_delta;
    A.fin = pl->fin;
    A.status = pl->status;
    A.iters = pl->iters;
    A.hist = hist_cap > 0 ? pl->hist : nullptr;
    A.hist_cap = hist_cap;
    A.max_iter = max_iter;
    A.tol = tol;
    A.quirk = pl->p.restart_quirk;
    A.x0_guard = pl->p.x0_guard;
    A.x0_rejected = pl->x0_rejected;
    A.dval = pl->dval;
    A.doff = pl->doff;
    A.dmax = pl->dia_on ? pl->dmax : 0;
    A.dpad = (long long)pl->total_chunks * kChunk;
    A.sync = pl->sync;
    CFM_HIP(hipMemsetAsync(pl->status, 0, sizeof(int) * pl->nsys, pl->stream));
    CFM_HIP(hipMemsetAsync(pl->x0_rejected, 0, sizeof(int) * pl->nsys, pl->stream));
    CFM_HIP(hipMemsetD32Async((hipDeviceptr_t)pl->fin, 0x7fffffff, pl->nsys, pl->stream));
    CFM_HIP(hipMemsetAsync(pl->iters, 0, sizeof(int) * pl->nsys, pl->stream));
    hipLaunchKernelGGL(reset_sync_kernel, dim3(8), dim3(256), 0, pl->stream, pl->sync, pl->spin_ticks, pl->grid);
    CFM_HIP(hipGetLastError());
    CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
    void *args[] = {&A};
    CFM_HIP(launch_persistent((const void *)batch_kernel(pl->sub), pl->grid, 256 * pl->sub, args, pl->stream, coop));
    CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
    CFM_HIP(hipEventSynchronize(pl->ev1));
    CFM_HIP(hipEventElapsedTime(ms, pl->ev0, pl->ev1));
    GridSync hs{};
    CFM_HIP(hipMemcpy(&hs, pl->sync, sizeof(hs), hipMemcpyDeviceToHost));
    if (hs.abort) return barrier_timeout("cg batch", *coop);
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_batch_solve(cfm_cg_batch_plan *pl, int *iters, int *status, double *hist, int hist_cap,
                                         cfm_timing *t) {
    if (!pl || hist_cap < 0) return CFM_EINVAL;
    const auto w0 = std::chrono::steady_clock::now();
    const int cap = hist ? hist_cap : 0;
    if (cap > pl->hist_cap) {
        if (pl->hist) (void)hipFree(pl->hist);
        pl->hist = nullptr;
        pl->hist_cap = 0;
        CFM_HIP(hipMalloc((void **)&pl->hist, sizeof(double) * (size_t)cap * pl->nsys));
        pl->hist_cap = cap;
    }
    float ms = 0.f;
    int coop = -1;
    cfm_status st = batch_run(pl, pl->p.max_iter, pl->p.tol, cap, &ms, &coop);
    if (st != CFM_OK) return st;
    std::vector<int> h_it(pl->nsys), h_st(pl->nsys);
    CFM_HIP(hipMemcpy(h_it.data(), pl->iters, sizeof(int) * pl->nsys, hipMemcpyDeviceToHost));
    CFM_HIP(hipMemcpy(h_st.data(), pl->status, sizeof(int) * pl->nsys, hipMemcpyDeviceToHost));
    if (hist && cap > 0)
        CFM_HIP(hipMemcpy(hist, pl->hist, sizeof(double) * (size_t)cap * pl->nsys, hipMemcpyDeviceToHost));
    cfm_status overall = CFM_OK;
    long long sum_it = 0;
    for (int sy = 0; sy < pl->nsys; ++sy) {
        const int code = h_st[sy] == kConverged ? CFM_OK
                         : h_st[sy] == kBreakdown ? CFM_EBREAKDOWN
                         : (h_st[sy] == kMaxIter && pl->p.tol > 0.0) ? CFM_ENOTCONV
                                                                    : CFM_OK;
        if (iters) iters[sy] = h_it[sy];
        if (status) status[sy] = code;
        if (code == CFM_EBREAKDOWN) overall = CFM_EBREAKDOWN;
        else if (code == CFM_ENOTCONV && overall == CFM_OK) overall = CFM_ENOTCONV;
        sum_it = std::max<long long>(sum_it, h_it[sy]);
    }
    if (t) {
        t->kernel_ms = ms;
        t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        t->grid_blocks = pl->grid;
        t->block_threads = 256 * pl->sub;
        t->iterations = (int)sum_it;
        t->outer_iterations = 0;
        t->cooperative = coop;
    }
    return overall;
}

// Autotuner.  Screen: time `trial_iters` fixed iterations of the batch for each
// workgroup shape (1, 2 or 4 sub-blocks of 256 threads), each grid size (1..occupancy
// workgroups per CU) and each storage the plan holds (SELL-64, and DIA when an auto
// plan's systems are banded).  Then time full solves (the plan's tol and max_iter,
// from the current x0) of the kTuneFinalists fastest and keep the fastest solve: the
// first iterations, where every system is active, do not show the tail, where the
// few systems still iterating leave most of the grid waiting at barriers.  With a
// cache file the winner is stored under a key of device, system count, row count and
// tolerance, and looked up first.
constexpr int kTuneFinalists = 3;

extern "C" cfm_status cfm_cg_batch_tune(cfm_cg_batch_plan *pl, int trial_iters, const char *cache_path,
                                        int *best_block, int *best_grid) {
    if (!pl || trial_iters < 1) return CFM_EINVAL;
    char key[512];
    std::snprintf(key, sizeof(key), "cg_batch|%s|cus=%d|nsys=%d|rows=%lld|chunks=%d|tol=%.3g", pl->dev.name,
                  pl->dev.cus, pl->nsys, pl->total_rows, pl->total_chunks, pl->p.tol);
    for (char *q = key; *q; ++q)  // the cache line is whitespace-separated
        if (*q == ' ' || *q == '\t') *q = '_';
    const bool have_sell = pl->p.matrix_format != 2, have_dia = pl->dmax > 0;
    auto use = [&](int sub, int grid, bool dia) {
        if ((dia && !have_dia) || (!dia && !have_sell)) return CFM_EINVAL;
        const cfm_status st = batch_configure(pl, sub, grid);
        if (st == CFM_OK) pl->dia_on = dia;
        return st;
    };
    auto report = [&]() {
        if (best_block) *best_block = 256 * pl->sub;
        if (best_grid) *best_grid = pl->grid;
    };
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char line[1024];
            while (std::fgets(line, sizeof(line), f)) {
                char k2[600];
                int s2 = 0, g2 = 0, d2 = 0;
                if (std::sscanf(line, "%599s %d %d %d", k2, &s2, &g2, &d2) == 4 && std::strcmp(k2, key) == 0 &&
                    use(s2, g2, d2 != 0) == CFM_OK) {
                    std::fclose(f);
                    report();
                    return CFM_OK;
                }
            }
            std::fclose(f);
        }
    }
    struct Cand {
        int sub, grid;
        bool dia;
        float ms;
    };
    std::vector<Cand> cands;
    const int subs[] = {1, 2, 4};
    for (int storage = 0; storage < 2; ++storage) {
        const bool dia = storage == 1;
        if ((dia && !have_dia) || (!dia && !have_sell)) continue;
        for (int sub : subs) {
            BatchKernelFn k = batch_kernel(sub);
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, 256 * sub, 0) != hipSuccess)
                continue;
            for (int m = 1; m <= per_cu; ++m) {
                if (use(sub, m * pl->dev.cus, dia) != CFM_OK) continue;
                float ms = 0.f;
                int coop = -1;
                cfm_status st = batch_run(pl, trial_iters, 0.0, 0, &ms, &coop);
                if (st != CFM_OK) return st;
                cands.push_back({pl->sub, pl->grid, dia, ms});
            }
        }
    }
    if (cands.empty()) return CFM_FAIL(CFM_EINVAL, "batched CG: no workgroup shape fits");
    std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    if (cands.size() > (size_t)kTuneFinalists) cands.resize(kTuneFinalists);
    Cand best = cands[0];
    float best_ms = 1e30f;
    for (const Cand &c : cands) {
        if (use(c.sub, c.grid, c.dia) != CFM_OK) continue;
        float ms = 0.f;
        int coop = -1;
        cfm_status st = batch_run(pl, pl->p.max_iter, pl->p.tol, 0, &ms, &coop);
        if (st != CFM_OK) return st;
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    cfm_status st = use(best.sub, best.grid, best.dia);
    if (st != CFM_OK) return st;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s %d %d %d\n", key, best.sub, best.grid, best.dia ? 1 : 0);
            std::fclose(f);
        }
    }
    report();
    return CFM_OK;
}

extern "C" int cfm_cg_batch_matrix_format(const cfm_cg_batch_plan *pl) {
    if (!pl) return -1;
    return pl->dia_on ? 2 : 1;
}

extern "C" cfm_status cfm_cg_batch_x0_rejected(cfm_cg_batch_plan *pl, int *flags) {
    if (!pl || !flags) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(flags, pl->x0_rejected, sizeof(int) * pl->nsys, hipMemcpyDeviceToHost));
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_batch_load_features(cfm_cg_batch_plan *pl, const float *features) {
    if (!pl || !features) return CFM_EINVAL;
    if (!pl->features) CFM_HIP(hipMalloc((void **)&pl->features, sizeof(float) * 2 * (size_t)pl->total_rows));
    CFM_HIP(hipMemcpy(pl->features, features, sizeof(float) * 2 * (size_t)pl->total_rows, hipMemcpyHostToDevice));
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_batch_warm_start(cfm_cg_batch_plan *pl, cfm_unet *net, double scale, cfm_timing *t) {
    if (!pl || !net || !(std::isfinite(scale))) return CFM_EINVAL;
    if (!pl->features) return CFM_EINVAL;
    const auto w0 = std::chrono::steady_clock::now();
    const UnetView nv = unet_view(net);
    const long long n3 = (long long)nv.n * nv.n * nv.n;
    for (int sy = 0; sy < pl->nsys; ++sy)
        if (pl->h_rows[sy] != n3) return CFM_EINVAL;  // every system must be the network's n^3 cube
    CFM_HIP(hipStreamSynchronize(pl->stream));
    CFM_HIP(hipEventRecord(pl->ev0, nv.stream));
    for (int g0 = 0; g0 < pl->nsys; g0 += nv.batch) {
        const int cnt = std::min(nv.batch, pl->nsys - g0);
        const float *src = pl->features + (size_t)g0 * 2 * n3;
        if (cnt < nv.batch) {  // a short last group runs from the network's own buffer, zero padded
            CFM_HIP(hipMemcpyAsync(nv.input, src, sizeof(float) * 2 * n3 * cnt, hipMemcpyDeviceToDevice, nv.stream));
            CFM_HIP(hipMemsetAsync(nv.input + 2 * n3 * cnt, 0, sizeof(float) * 2 * n3 * (nv.batch - cnt), nv.stream));
            src = nv.input;
        }
        const float *pred = nullptr;
        cfm_status st = unet_forward_device(net, src, &pred);
        if (st != CFM_OK) return st;
        const size_t count = (size_t)cnt * n3;
        hipLaunchKernelGGL(scatter_x0_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, nv.stream, pred,
                           pl->sys_chunk0, g0, cnt, (int)n3, scale, pl->x0);
        CFM_HIP(hipGetLastError());
    }
    CFM_HIP(hipEventRecord(pl->ev1, nv.stream));
    CFM_HIP(hipEventSynchronize(pl->ev1));
    float ms = 0.f;
    CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
    if (t) {
        t->kernel_ms = ms;
        t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        t->grid_blocks = 0;
        t->block_threads = 0;
        t->iterations = (pl->nsys + nv.batch - 1) / nv.batch;
        t->outer_iterations = 0;
        t->cooperative = -1;
    }
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_batch_smooth_x0(cfm_cg_batch_plan *pl, int sweeps, double omega, cfm_timing *t) {
    if (!pl || sweeps < 0 || !(omega > 0.0 && omega < 2.0)) return CFM_EINVAL;
    const auto w0 = std::chrono::steady_clock::now();
    const int nrows = pl->total_chunks * kChunk;
    const Sell A{pl->slice_ptr, pl->scol, pl->sval};
    CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
    double *src = pl->x0, *dst = pl->r;  // r is scratch until the next solve initialises it
    for (int k = 0; k < sweeps; ++k) {
        if (pl->dia_on)
            hipLaunchKernelGGL(jacobi_smooth_dia_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0,
                               pl->stream, pl->dval, pl->doff, pl->dmax, (long long)nrows, pl->sys_chunk0,
                               pl->sys_rows, pl->nsys, pl->b, src, dst, nrows, omega);
        else
            hipLaunchKernelGGL(jacobi_smooth_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, pl->stream,
                               A, pl->b, src, dst, nrows, omega);
        CFM_HIP(hipGetLastError());
        std::swap(src, dst);
    }
    if (src != pl->x0)
        CFM_HIP(hipMemcpyAsync(pl->x0, src, sizeof(double) * nrows, hipMemcpyDeviceToDevice, pl->stream));
    CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
    CFM_HIP(hipEventSynchronize(pl->ev1));
    float ms = 0.f;
    CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
    if (t) {
        t->kernel_ms = ms;
        t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        t->grid_blocks = (nrows + 255) / 256;
        t->block_threads = 256;
        t->iterations = sweeps;
        t->outer_iterations = 0;
        t->cooperative = -1;
    }
    return CFM_OK;
}
