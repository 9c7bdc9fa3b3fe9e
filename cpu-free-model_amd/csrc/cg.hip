// CPU-free Conjugate Gradient for gfx950: Standard, Pipelined and an FP32
// inner CG inside an FP64 iterative-refinement loop, each as one cooperative
// persistent kernel; plus a host-driven ("discrete") schedule of the same
// device phases for comparison.
//
// Reference algorithms restated:
//   Standard   CG/src/common.cu:399-538 and CG/src/single-stream/standard.cu:158-299
//   Pipelined  CG/src/common.cu:609-780 and CG/src/single-stream/pipelined.cu:176-318
//   SpMV       CG/src/common.cu:150-168 (per-row sum in stored column order)
//   dot        CG/src/common.cu:357-389 (double accumulation)
//
// MI355X design:
//   * The CSR matrix is stored as SELL-64: slices of 64 consecutive rows, the
//     j-th entry of every row of a slice contiguous, so the 64 lanes of a
//     wavefront read col/val with one coalesced load each.  Per-row sums run
//     in the stored column order, so results equal the CSR formula exactly.
//     Padding entries carry col = -1 and are skipped.
//   * Rows are grouped in chunks of 1024 (16 slices); workgroup b owns chunks
//     b, b+G, ... in every phase, so the vectors a workgroup updates stay in
//     its XCD's L2 between phases.
//   * Dot products are deterministic and independent of the grid size: each
//     chunk produces one partial with a fixed tree, and every workgroup
//     reduces all partials in the same fixed order after the barrier (no
//     atomics, no second barrier).  The CPU oracle replicates the tree, so
//     residual histories agree bit for bit.
//   * Pipelined CG needs one device-wide barrier per iteration: the two dot
//     products, q = A w and all six vector updates are fused, with w and
//     the partials double-buffered.  Standard CG needs three.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "cg_common.hpp"

namespace cfm {

struct CgArgs {
    int n;
    int nchunks;
    Sell A64;  // FP64 values
    Sell A32;  // FP32 values (IR only)
    const double *b;
    const double *x0;
    double *x;
    double *v[6];        // work vectors (meaning depends on the method)
    float *f[4];         // FP32 work vectors (IR)
    double *part;        // partial sums, [8][nchunks]
    double *hist;
    int hist_cap;
    int max_iter;
    double tol;
    int quirk;
    int max_outer;
    double inner_tol;
    int *info;           // [0] status, [1] iterations, [2] outer iterations
    GridSync *sync;
    unsigned long long *timers;  // phase-timer build: [workgroup][kTimerSlots] ticks
};


// Iterate the rows this thread owns: sub-block `sub_` of workgroup b takes
// chunk c = b*SUB + sub_ (+ multiples of gridDim.x*SUB); inside a chunk its
// thread t owns rows c*1024 + i*256 + t, i = 0..3 (slice c*16 + i*4 + t/64).
// The loop trip count is uniform over the workgroup (inactive sub-blocks
// still take part in the __syncthreads of the chunk sums).
#define CFM_FOR_MY_ROWS(...)                                                               \
    const int sub_ = threadIdx.x >> 8;                                                     \
    const int t_ = threadIdx.x & (kThreads - 1);                                           \
    for (int base_ = blockIdx.x * SUB; base_ < A.nchunks; base_ += gridDim.x * SUB) {      \
        const int c_ = base_ + sub_;                                                       \
        const bool cact_ = c_ < A.nchunks;                                                 \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) {                                  \
            const int slice = c_ * kSlicesPerChunk + i_ * 4 + (t_ >> 6);                   \
            const int row = c_ * kChunk + i_ * kThreads + t_;                               \
            const int lane = t_ & 63;                                                       \
            (void)slice;                                                                    \
            (void)lane;                                                                     \
            if (cact_ && row < A.n) {                                                       \
                __VA_ARGS__                                                                 \
            }                                                                               \
        }                                                                                   \
        CHUNK_END                                                                           \
    }

// Store that writes through the XCD's L2 (a relaxed agent-scope atomic store:
// a vector store with scope bit sc1 on gfx950) instead of leaving a dirty
// line there for the next device barrier's release to write back.  Measured
// on the 128^3 Poisson system: Standard CG (s, x, r, p; three barriers per
// iteration) 82.1 -> 76.4 us per iteration; the refinement's FP32 inner
// iteration (s, d, r, p) 60.6 -> 59.7 us; Pipelined CG (one barrier per
// iteration) gains only for w, the vector the next SpMV gathers (71.6 -> 70.7
// us), and runs slower with its other vectors written through.  Being a
// compiler-visible store, it is counted in the compiler's own wait counters.
// Ordering against other workgroups comes only from the device barrier (a
// workgroup barrier, then one thread's agent-scope release, which the HIP
// memory model makes cumulative over the workgroup's earlier stores; an
// agent-scope acquire on exit); the relaxed atomic adds no ordering of its own.
__device__ __forceinline__ void store_wt(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_wt(float *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// =============================================================== phases ===
// Standard CG.  v[0]=r, v[1]=p, v[2]=s.  part: [0]=gamma, [1]=delta, [2]=bb.
struct StdPhases {
    // r = b - A x0 ; p = r ; x = x0 ; partials <r,r>, <b,b>
    template <int SUB>
    static __device__ void init(const CgArgs &A, double *lds) {
        double *r = A.v[0], *p = A.v[1];
        double ag = 0.0, ab = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sg = sub_sum<SUB>(ag, lds);                                           \
        const double sb = sub_sum<SUB>(ab, lds);                                           \
        if (t_ == 0 && cact_) {                                                             \
            A.part[0 * A.nchunks + c_] = sg;                                                \
            A.part[2 * A.nchunks + c_] = sb;                                                \
        }                                                                                   \
        ag = 0.0;                                                                           \
        ab = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double ax = sell_row<double, double>(A.A64, A.x0, slice, lane);
            const double bi = A.b[row];
            const double ri = fma(-1.0, ax, bi);
            r[row] = ri;
            p[row] = ri;
            A.x[row] = A.x0[row];
            ag = fma(ri, ri, ag);
            ab = fma(bi, bi, ab);
        })
#undef CHUNK_END
    }
    // s = A p ; partial <p,s>
    template <int SUB>
    static __device__ void spmv_dot(const CgArgs &A, double *lds) {
        const double *p = A.v[1];
        double *s = A.v[2];
        double ad = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sd = sub_sum<SUB>(ad, lds);                                           \
        if (t_ == 0 && cact_) A.part[1 * A.nchunks + c_] = sd;                              \
        ad = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double si = sell_row<double, double>(A.A64, p, slice, lane);
            store_wt(&s[row], si);
            ad = fma(p[row], si, ad);
        })
#undef CHUNK_END
    }
    // x += alpha p ; r -= alpha s ; partial <r,r>
    template <int SUB>
    static __device__ void update_xr(const CgArgs &A, double alpha, double *lds) {
        double *r = A.v[0];
        const double *p = A.v[1], *s = A.v[2];
        double ag = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sg = sub_sum<SUB>(ag, lds);                                           \
        if (t_ == 0 && cact_) A.part[0 * A.nchunks + c_] = sg;                              \
        ag = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            store_wt(&A.x[row], fma(alpha, p[row], A.x[row]));
            const double ri = fma(-alpha, s[row], r[row]);
            store_wt(&r[row], ri);
            ag = fma(ri, ri, ag);
        })
#undef CHUNK_END
    }
    // p = r + beta p
    template <int SUB>
    static __device__ void update_p(const CgArgs &A, double beta) {
        const double *r = A.v[0];
        double *p = A.v[1];
#define CHUNK_END
        CFM_FOR_MY_ROWS({ store_wt(&p[row], fma(beta, p[row], r[row])); })
#undef CHUNK_END
    }
};

// Pipelined CG.  v[0]=r, v[1]=z, v[2]=s, v[3]=p, v[4]=w0, v[5]=w1.
// part: [0,1] = delta (buffers 0/1), [2,3] = gamma (buffers 0/1), [4] = bb.
struct PipePhases {
    // r = b - A x0 ; x = x0 ; z = s = p = 0 ; partial <b,b>
    template <int SUB>
    static __device__ void init_r(const CgArgs &A, double *lds) {
        double ab = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sb = sub_sum<SUB>(ab, lds);                                           \
        if (t_ == 0 && cact_) A.part[4 * A.nchunks + c_] = sb;                              \
        ab = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double ax = sell_row<double, double>(A.A64, A.x0, slice, lane);
            const double bi = A.b[row];
            A.v[0][row] = fma(-1.0, ax, bi);
            A.x[row] = A.x0[row];
            A.v[1][row] = 0.0;
            A.v[2][row] = 0.0;
            A.v[3][row] = 0.0;
            ab = fma(bi, bi, ab);
        })
#undef CHUNK_END
    }
    // w0 = A r ; partials <r,r>, <r,w0> into buffer 0
    template <int SUB>
    static __device__ void init_w(const CgArgs &A, double *lds) {
        double ad = 0.0, ag = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sd = sub_sum<SUB>(ad, lds);                                           \
        const double sg = sub_sum<SUB>(ag, lds);                                           \
        if (t_ == 0 && cact_) {                                                             \
            A.part[0 * A.nchunks + c_] = sd;                                                \
            A.part[2 * A.nchunks + c_] = sg;                                                \
        }                                                                                   \
        ad = 0.0;                                                                           \
        ag = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double wi = sell_row<double, double>(A.A64, A.v[0], slice, lane);
            A.v[4][row] = wi;
            const double ri = A.v[0][row];
            ad = fma(ri, ri, ad);
            ag = fma(ri, wi, ag);
        })
#undef CHUNK_END
    }
    // q = A w_cur ; z,s,p,x,r,w updates ; partials for the next iteration
    template <int SUB>
    static __device__ void iterate(const CgArgs &A, int buf, double alpha, double beta,
                                   double *lds) {
        const double *wc = A.v[4 + buf];
        double *wn = A.v[5 - buf];
        double *r = A.v[0], *z = A.v[1], *s = A.v[2], *p = A.v[3];
        const int nb = buf ^ 1;
        double ad = 0.0, ag = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sd = sub_sum<SUB>(ad, lds);                                           \
        const double sg = sub_sum<SUB>(ag, lds);                                           \
        if (t_ == 0 && cact_) {                                                             \
            A.part[(0 + nb) * A.nchunks + c_] = sd;                                         \
            A.part[(2 + nb) * A.nchunks + c_] = sg;                                         \
        }                                                                                   \
        ad = 0.0;                                                                           \
        ag = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double qi = sell_row<double, double>(A.A64, wc, slice, lane);
            const double wi = wc[row];
            const double zi = fma(beta, z[row], qi);
            const double si = fma(beta, s[row], wi);
            const double pi = fma(beta, p[row], r[row]);
            A.x[row] = fma(alpha, pi, A.x[row]);
            const double ri = fma(-alpha, si, r[row]);
            const double wn_i = fma(-alpha, zi, wi);
            z[row] = zi;
            s[row] = si;
            p[row] = pi;
            r[row] = ri;
            store_wt(&wn[row], wn_i);  // gathered by other workgroups next iteration
            ad = fma(ri, ri, ad);
            ag = fma(ri, wn_i, ag);
        })
#undef CHUNK_END
    }
};

// FP32 inner CG inside FP64 refinement.  v[0] = r64.  f[0]=rf, f[1]=p, f[2]=s, f[3]=d.
// part: [0]=rr (outer), [1]=g (inner <r,r>), [2]=delta, [3]=bb.
struct IrPhases {
    // r64 = b - A x ; rf = (float) r64 ; p = rf ; d = 0 ; partials rr, g (and bb when first)
    template <int SUB>
    static __device__ void outer(const CgArgs &A, bool first, double *lds) {
        double arr = 0.0, ag = 0.0, ab = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double s1 = sub_sum<SUB>(arr, lds);                                          \
        const double s2 = sub_sum<SUB>(ag, lds);                                           \
        const double s3 = first ? sub_sum<SUB>(ab, lds) : 0.0;                             \
        if (t_ == 0 && cact_) {                                                             \
            A.part[0 * A.nchunks + c_] = s1;                                                \
            A.part[1 * A.nchunks + c_] = s2;                                                \
            if (first) A.part[3 * A.nchunks + c_] = s3;                                     \
        }                                                                                   \
        arr = 0.0;                                                                          \
        ag = 0.0;                                                                           \
        ab = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const double *xs = first ? A.x0 : A.x;
            const double ax = sell_row<double, double>(A.A64, xs, slice, lane);
            const double bi = A.b[row];
            const double ri = fma(-1.0, ax, bi);
            if (first) A.x[row] = A.x0[row];
            A.v[0][row] = ri;
            const float rf = (float)ri;
            A.f[0][row] = rf;
            A.f[1][row] = rf;
            A.f[3][row] = 0.0f;
            arr = fma(ri, ri, arr);
            ag = fma((double)rf, (double)rf, ag);
            ab = fma(bi, bi, ab);
        })
#undef CHUNK_END
    }
    template <int SUB>
    static __device__ void spmv_dot(const CgArgs &A, double *lds) {
        const float *p = A.f[1];
        float *s = A.f[2];
        double ad = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sd = sub_sum<SUB>(ad, lds);                                           \
        if (t_ == 0 && cact_) A.part[2 * A.nchunks + c_] = sd;                              \
        ad = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            const float si = sell_row<float, float>(A.A32, p, slice, lane);
            store_wt(&s[row], si);
            ad = fma((double)p[row], (double)si, ad);
        })
#undef CHUNK_END
    }
    template <int SUB>
    static __device__ void update_dr(const CgArgs &A, float af, double *lds) {
        float *rf = A.f[0], *d = A.f[3];
        const float *p = A.f[1], *s = A.f[2];
        double ag = 0.0;
#define CHUNK_END                                                                           \
    {                                                                                       \
        const double sg = sub_sum<SUB>(ag, lds);                                           \
        if (t_ == 0 && cact_) A.part[1 * A.nchunks + c_] = sg;                              \
        ag = 0.0;                                                                           \
    }
        CFM_FOR_MY_ROWS({
            store_wt(&d[row], fmaf(af, p[row], d[row]));
            const float ri = fmaf(-af, s[row], rf[row]);
            store_wt(&rf[row], ri);
            ag = fma((double)ri, (double)ri, ag);
        })
#undef CHUNK_END
    }
    template <int SUB>
    static __device__ void update_p(const CgArgs &A, float bf) {
        const float *rf = A.f[0];
        float *p = A.f[1];
#define CHUNK_END
        CFM_FOR_MY_ROWS({ store_wt(&p[row], fmaf(bf, p[row], rf[row])); })
#undef CHUNK_END
    }
    template <int SUB>
    static __device__ void update_x(const CgArgs &A) {
        const float *d = A.f[3];
#define CHUNK_END
        CFM_FOR_MY_ROWS({ A.x[row] = A.x[row] + (double)d[row]; })
#undef CHUNK_END
    }
};

// ======================================================= persistent kernels
__device__ __forceinline__ void write_info(const CgArgs &A, int status, int iters, int outer) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.info[0] = status;
        A.info[1] = iters;
        A.info[2] = outer;
    }
}
__device__ __forceinline__ void write_hist(const CgArgs &A, int k, double v) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && A.hist && k < A.hist_cap) A.hist[k] = v;
}

// The persistent kernels.  TIMED builds a second copy that records every
// workgroup's time per phase (PhaseClock); the default build has no timer code.
template <int SUB, bool TIMED>
__global__ void __launch_bounds__(256 * SUB) cg_standard_kernel(CgArgs A) {
    __shared__ double lds[4 * SUB];
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    unsigned int epoch = 0;
    StdPhases::init<SUB>(A, lds);
    clk.mark(CFM_PHASE_OTHER);
    if (!grid_barrier(A.sync, epoch)) return;
    clk.mark(CFM_PHASE_BARRIER);
    double gamma = reduce_partials<SUB>(A.part + 0 * A.nchunks, A.nchunks, lds);
    const double bb = reduce_partials<SUB>(A.part + 2 * A.nchunks, A.nchunks, lds);
    const double thr = A.tol > 0.0 ? A.tol * A.tol * bb : 0.0;
    write_hist(A, 0, gamma);
    clk.mark(CFM_PHASE_REDUCE);
    int k = 0, status = CFM_OK;
    bool converged = gamma <= thr;
    while (!converged && k < A.max_iter) {
        StdPhases::spmv_dot<SUB>(A, lds);
        clk.mark(CFM_PHASE_SPMV);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        const double delta = reduce_partials<SUB>(A.part + 1 * A.nchunks, A.nchunks, lds);
        clk.mark(CFM_PHASE_REDUCE);
        if (!(delta > 0.0) || !is_finite(delta)) {
            status = CFM_EBREAKDOWN;
            break;
        }
        const double alpha = gamma / delta;
        StdPhases::update_xr<SUB>(A, alpha, lds);
        clk.mark(CFM_PHASE_UPDATE);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        const double gamma_new = reduce_partials<SUB>(A.part + 0 * A.nchunks, A.nchunks, lds);
        ++k;
        write_hist(A, k, gamma_new);
        clk.mark(CFM_PHASE_REDUCE);
        if (gamma_new <= thr) {
            converged = true;
            break;
        }
        const double beta = gamma_new / gamma;
        StdPhases::update_p<SUB>(A, beta);
        clk.mark(CFM_PHASE_UPDATE);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        gamma = gamma_new;
    }
    if (status == CFM_OK && !converged && A.tol > 0.0) status = CFM_ENOTCONV;
    write_info(A, status, k, 0);
    clk.flush(A.timers);
}

template <int SUB, bool TIMED>
__global__ void __launch_bounds__(256 * SUB) cg_pipelined_kernel(CgArgs A) {
    __shared__ double lds[4 * SUB];
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    unsigned int epoch = 0;
    PipePhases::init_r<SUB>(A, lds);
    clk.mark(CFM_PHASE_OTHER);
    if (!grid_barrier(A.sync, epoch)) return;
    clk.mark(CFM_PHASE_BARRIER);
    PipePhases::init_w<SUB>(A, lds);
    clk.mark(CFM_PHASE_OTHER);
    if (!grid_barrier(A.sync, epoch)) return;
    clk.mark(CFM_PHASE_BARRIER);
    const double bb = reduce_partials<SUB>(A.part + 4 * A.nchunks, A.nchunks, lds);
    const double thr = A.tol > 0.0 ? A.tol * A.tol * bb : 0.0;
    double alpha = 0.0, beta = 0.0, delta_prev = 0.0;
    int k = 0, status = CFM_OK;
    bool converged = false;
    for (;;) {
        const int buf = k & 1;
        const double delta = reduce_partials<SUB>(A.part + (0 + buf) * A.nchunks, A.nchunks, lds);
        const double gam = reduce_partials<SUB>(A.part + (2 + buf) * A.nchunks, A.nchunks, lds);
        write_hist(A, k, delta);
        clk.mark(CFM_PHASE_REDUCE);
        if (delta <= thr) {
            converged = true;
            break;
        }
        if (k >= A.max_iter) break;
        double denom;
        if (A.quirk ? (k > 1) : (k > 0)) {
            beta = delta / delta_prev;
            denom = gam - (beta / alpha) * delta;
        } else {
            beta = 0.0;
            denom = gam;
        }
        if (!(denom > 0.0) || !is_finite(denom)) {
            status = CFM_EBREAKDOWN;
            break;
        }
        alpha = delta / denom;
        PipePhases::iterate<SUB>(A, buf, alpha, beta, lds);
        clk.mark(CFM_PHASE_SPMV);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        delta_prev = delta;
        ++k;
    }
    if (status == CFM_OK && !converged && A.tol > 0.0) status = CFM_ENOTCONV;
    write_info(A, status, k, 0);
    clk.flush(A.timers);
}

template <int SUB, bool TIMED>
__global__ void __launch_bounds__(256 * SUB) cg_ir_kernel(CgArgs A) {
    __shared__ double lds[4 * SUB];
    __shared__ unsigned long long clk_lds[kClockWords];
    PhaseClock<TIMED> clk{clk_lds};
    clk.start();
    unsigned int epoch = 0;
    double thr = 0.0;
    int outer = 0, inner_total = 0, status = CFM_ENOTCONV;
    for (;;) {
        const bool first = outer == 0;
        IrPhases::outer<SUB>(A, first, lds);
        clk.mark(CFM_PHASE_OTHER);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        const double rr = reduce_partials<SUB>(A.part + 0 * A.nchunks, A.nchunks, lds);
        if (first) {
            const double bb = reduce_partials<SUB>(A.part + 3 * A.nchunks, A.nchunks, lds);
            thr = A.tol * A.tol * bb;
        }
        write_hist(A, outer, rr);
        clk.mark(CFM_PHASE_REDUCE);
        if (rr <= thr) {
            status = CFM_OK;
            break;
        }
        if (outer == A.max_outer) break;
        double g = reduce_partials<SUB>(A.part + 1 * A.nchunks, A.nchunks, lds);
        // no refinement is solved beyond half the outer target
        const double thr_in = fmax(A.inner_tol * A.inner_tol * g, 0.25 * thr);
        clk.mark(CFM_PHASE_REDUCE);
        int j = 0;
        bool brk = false;
        while (j < A.max_iter) {
            IrPhases::spmv_dot<SUB>(A, lds);
            clk.mark(CFM_PHASE_SPMV);
            if (!grid_barrier(A.sync, epoch)) return;
            clk.mark(CFM_PHASE_BARRIER);
            const double delta = reduce_partials<SUB>(A.part + 2 * A.nchunks, A.nchunks, lds);
            clk.mark(CFM_PHASE_REDUCE);
            if (!(delta > 0.0) || !is_finite(delta)) {
                brk = true;
                break;
            }
            const float af = (float)(g / delta);
            IrPhases::update_dr<SUB>(A, af, lds);
            clk.mark(CFM_PHASE_UPDATE);
            if (!grid_barrier(A.sync, epoch)) return;
            clk.mark(CFM_PHASE_BARRIER);
            const double gn = reduce_partials<SUB>(A.part + 1 * A.nchunks, A.nchunks, lds);
            clk.mark(CFM_PHASE_REDUCE);
            ++j;
            if (gn <= thr_in) break;
            const float bf = (float)(gn / g);
            IrPhases::update_p<SUB>(A, bf);
            clk.mark(CFM_PHASE_UPDATE);
            if (!grid_barrier(A.sync, epoch)) return;
            clk.mark(CFM_PHASE_BARRIER);
            g = gn;
        }
        inner_total += j;
        IrPhases::update_x<SUB>(A);
        clk.mark(CFM_PHASE_UPDATE);
        if (!grid_barrier(A.sync, epoch)) return;
        clk.mark(CFM_PHASE_BARRIER);
        ++outer;
        if (brk && j == 0) {
            status = CFM_EBREAKDOWN;
            break;
        }
    }
    write_info(A, status, inner_total, outer);
    clk.flush(A.timers);
}

// ===================================================== discrete kernels ===
// One launch per phase; scalars are reduced by a one-workgroup kernel and
// read back by the host, which forms alpha and beta itself.
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_std_init(CgArgs A) {
    __shared__ double lds[4 * SUB];
    StdPhases::init<SUB>(A, lds);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_std_spmv_dot(CgArgs A) {
    __shared__ double lds[4 * SUB];
    StdPhases::spmv_dot<SUB>(A, lds);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_std_update_xr(CgArgs A, double alpha) {
    __shared__ double lds[4 * SUB];
    StdPhases::update_xr<SUB>(A, alpha, lds);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_std_update_p(CgArgs A, double beta) {
    StdPhases::update_p<SUB>(A, beta);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_pipe_init_r(CgArgs A) {
    __shared__ double lds[4 * SUB];
    PipePhases::init_r<SUB>(A, lds);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_pipe_init_w(CgArgs A) {
    __shared__ double lds[4 * SUB];
    PipePhases::init_w<SUB>(A, lds);
}
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_pipe_iterate(CgArgs A, int buf, double alpha, double beta) {
    __shared__ double lds[4 * SUB];
    PipePhases::iterate<SUB>(A, buf, alpha, beta, lds);
}
// out[i] = reduce(part + idx[i]*nchunks), i < count
template <int SUB>
__global__ void __launch_bounds__(256 * SUB) k_reduce(CgArgs A, int i0, int i1, double *out) {
    __shared__ double lds[4 * SUB];
    const double a = reduce_partials<SUB>(A.part + i0 * A.nchunks, A.nchunks, lds);
    double b = 0.0;
    if (i1 >= 0) b = reduce_partials<SUB>(A.part + i1 * A.nchunks, A.nchunks, lds);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

}  // namespace cfm

using namespace cfm;

using CgKernelFn = void (*)(CgArgs);

// The kernels of one workgroup shape (SUB sub-blocks of 256 threads).
struct CgKernels {
    CgKernelFn standard, pipelined, ir;
    CgKernelFn standard_timed, pipelined_timed, ir_timed;
    void (*std_init)(CgArgs);
    void (*std_spmv_dot)(CgArgs);
    void (*std_update_xr)(CgArgs, double);
    void (*std_update_p)(CgArgs, double);
    void (*pipe_init_r)(CgArgs);
    void (*pipe_init_w)(CgArgs);
    void (*pipe_iterate)(CgArgs, int, double, double);
    void (*reduce)(CgArgs, int, int, double *);
    CgKernelFn persistent(int method, bool timed) const {
        if (timed)
            return method == CFM_CG_STANDARD ? standard_timed : method == CFM_CG_PIPELINED ? pipelined_timed : ir_timed;
        return method == CFM_CG_STANDARD ? standard : method == CFM_CG_PIPELINED ? pipelined : ir;
    }
};

template <int SUB>
static CgKernels kernels_for() {
    return CgKernels{cg_standard_kernel<SUB, false>, cg_pipelined_kernel<SUB, false>, cg_ir_kernel<SUB, false>,
                     cg_standard_kernel<SUB, true>,  cg_pipelined_kernel<SUB, true>,  cg_ir_kernel<SUB, true>,
                     k_std_init<SUB>,                k_std_spmv_dot<SUB>,             k_std_update_xr<SUB>,
                     k_std_update_p<SUB>,            k_pipe_init_r<SUB>,              k_pipe_init_w<SUB>,
                     k_pipe_iterate<SUB>,            k_reduce<SUB>};
}

static bool kernel_table(int sub, CgKernels *out) {
    switch (sub) {
        case 1: *out = kernels_for<1>(); return true;
        case 2: *out = kernels_for<2>(); return true;
        case 3: *out = kernels_for<3>(); return true;
        case 4: *out = kernels_for<4>(); return true;
        default: return false;
    }
}

template <typename... Args>
static hipError_t launch(void (*f)(Args...), dim3 grid, dim3 block, hipStream_t s, Args... args) {
    void *ptrs[] = {&args...};
    return hipLaunchKernel((const void *)f, grid, block, ptrs, 0, s);
}

// ==================================================================== host
struct cfm_cg_plan {
    cfm_cg_params p{};
    DeviceProps dev{};
    int n = 0, nnz = 0, nslices = 0, nchunks = 0;
    int *slice_ptr = nullptr, *scol = nullptr;
    double *sval = nullptr;
    float *sval32 = nullptr;
    int *doff = nullptr;  // DIA offsets (ndiag > 0)
    int ndiag = 0;
    double *b = nullptr, *x0 = nullptr, *x = nullptr;
    double *v[6] = {};
    float *f[4] = {};
    double *part = nullptr, *hist = nullptr, *scal = nullptr, *scal_h = nullptr;
    int *info = nullptr;
    int hist_cap = 0;
    GridSync *sync = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int grid = 0;
    int sub = 4;  // 256-thread sub-blocks per workgroup
    CgKernels K{};
    unsigned long long spin_ticks = 0;
    double clock_khz = 100000.0;           // s_memrealtime rate
    unsigned long long *timers = nullptr;  // phase_timers: [grid][kTimerSlots] ticks
    cfm_phase_times phases{};              // breakdown of the last solve
};

// Build SELL-64 on the host from CSR (rows padded to the slice width, col -1).
static cfm_status build_sell(int n, const int *rowptr, const int *col, const double *val, int base,
                             std::vector<int> &sp, std::vector<int> &sc, std::vector<double> &sv) {
    const int nslices = (n + 63) / 64;
    sp.assign(nslices + 1, 0);
    long long total = 0;
    for (int s = 0; s < nslices; ++s) {
        int w = 0;
        for (int r = s * 64; r < std::min(n, s * 64 + 64); ++r) {
            const int len = rowptr[r + 1] - rowptr[r];
            if (len < 0) return CFM_EINVAL;
            w = std::max(w, len);
        }
        sp[s] = (int)total;
        total += 64LL * w;
        if (total > 0x7fffffffLL) return CFM_EINVAL;
    }
    sp[nslices] = (int)total;
    sc.assign((size_t)total, -1);
    sv.assign((size_t)total, 0.0);
    for (int r = 0; r < n; ++r) {
        const int s = r / 64, lane = r % 64;
        const int beg = rowptr[r] - base;
        const int len = rowptr[r + 1] - rowptr[r];
        for (int j = 0; j < len; ++j) {
            const int c = col[beg + j] - base;
            if (c < 0 || c >= n) return CFM_EINVAL;
            sc[(size_t)sp[s] + (size_t)j * 64 + lane] = c;
            sv[(size_t)sp[s] + (size_t)j * 64 + lane] = val[beg + j];
        }
    }
    return CFM_OK;
}

// DIA storage when the CSR is banded: at most kMaxDiag distinct column
// offsets, rows sorted by column, and the zero fill at most half the nnz.
// Returns false (SELL-64 is used) otherwise.
constexpr int kMaxDiag = 16;
static bool build_dia(int n, int npad, const int *rowptr, const int *col, const double *val, int base,
                      std::vector<int> &offs, std::vector<double> &dv) {
    offs.clear();
    for (int r = 0; r < n; ++r) {
        int prev = -1;
        for (int j = rowptr[r] - base; j < rowptr[r + 1] - base; ++j) {
            const int c = col[j] - base;
            if (c <= prev) return false;  // unsorted or duplicate columns
            prev = c;
            const int o = c - r;
            if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
                if ((int)offs.size() == kMaxDiag) return false;
                offs.push_back(o);
            }
        }
    }
    if (offs.empty()) return false;
    const long long nnz = rowptr[n] - rowptr[0];
    if ((long long)offs.size() * n > nnz + nnz / 2) return false;
    std::sort(offs.begin(), offs.end());
    dv.assign(offs.size() * (size_t)npad, 0.0);
    for (int r = 0; r < n; ++r)
        for (int j = rowptr[r] - base; j < rowptr[r + 1] - base; ++j) {
            const int o = col[j] - base - r;
            const size_t d = std::lower_bound(offs.begin(), offs.end(), o) - offs.begin();
            dv[d * npad + r] = val[j];
        }
    return true;
}

extern "C" void cfm_cg_plan_destroy(cfm_cg_plan *pl) {
    if (!pl) return;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    void *ptrs[] = {pl->slice_ptr, pl->scol, pl->sval, pl->sval32, pl->doff, pl->b, pl->x0, pl->x,
                    pl->part, pl->hist, pl->scal, pl->info, pl->sync, pl->timers};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (double *q : pl->v)
        if (q) (void)hipFree(q);
    for (float *q : pl->f)
        if (q) (void)hipFree(q);
    if (pl->scal_h) (void)hipHostFree(pl->scal_h);
    if (pl->ev0) (void)hipEventDestroy(pl->ev0);
    if (pl->ev1) (void)hipEventDestroy(pl->ev1);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

extern "C" cfm_status cfm_cg_plan_create(int n, int nnz, const int *rowptr, const int *col,
                                         const double *val, int index_base, int csr_on_device,
                                         const cfm_cg_params *p, cfm_cg_plan **out) {
    if (!out) return CFM_EINVAL;
    *out = nullptr;
    if (!p || n < 1 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) return CFM_EINVAL;
    if (index_base != 0 && index_base != 1) return CFM_EINVAL;
    if (p->method < CFM_CG_STANDARD || p->method > CFM_CG_IR_F32) return CFM_EINVAL;
    if (p->mode != CFM_CG_PERSISTENT && p->mode != CFM_CG_DISCRETE) return CFM_EINVAL;
    if (p->mode == CFM_CG_DISCRETE && p->method == CFM_CG_IR_F32) return CFM_ENOTSUP;
    if (p->max_iter < 0 || n > 0x7fffffff - 2048) return CFM_EINVAL;
    if (p->matrix_format < 0 || p->matrix_format > 2) return CFM_EINVAL;
    // a host CSR is checked before any device work; a borrowed device CSR right after its copy
    if (!csr_on_device) {
        const cfm_status vst = check_csr(n, nnz, rowptr, col, index_base, n, "cg");
        if (vst != CFM_OK) return vst;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CFM_EDEVICE;

    // host copies of the CSR (device CSR is borrowed and read back once)
    std::vector<int> h_rowptr, h_col;
    std::vector<double> h_val;
    const int *rp = rowptr, *cp = col;
    const double *vp = val;
    if (csr_on_device) {
        h_rowptr.resize((size_t)n + 1);
        h_col.resize((size_t)nnz);
        h_val.resize((size_t)nnz);
        if (hipMemcpy(h_rowptr.data(), rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost) != hipSuccess ||
            (nnz > 0 && hipMemcpy(h_col.data(), col, sizeof(int) * nnz, hipMemcpyDeviceToHost) != hipSuccess) ||
            (nnz > 0 && hipMemcpy(h_val.data(), val, sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess))
            return CFM_EDEVICE;
        rp = h_rowptr.data();
        cp = h_col.data();
        vp = h_val.data();
    }
    if (csr_on_device) {  // every builder below (DIA, SELL-64) reads the CSR only after this check
        const cfm_status vst = check_csr(n, nnz, rp, cp, index_base, n, "cg");
        if (vst != CFM_OK) return vst;
    }
    std::vector<int> sp, sc, doffs;
    std::vector<double> sv;
    const int npad = (n + kChunk - 1) / kChunk * kChunk;
    // auto: DIA for Pipelined CG on large systems (measured 70 vs 80 us per
    // iteration at 128^3 = 2.1M rows, where the index-free matrix and the vectors
    // fit the MALL); SELL-64 for small systems (64^3: 27.8 vs 22.2 us per
    // iteration, latency-bound, where the 8-wide SELL loop wins), for Standard CG
    // and for the FP32 refinement
    constexpr int kDiaMinRows = 1 << 20;
    bool dia = false;
    if (p->matrix_format == 2 ||
        (p->matrix_format == 0 && p->method == CFM_CG_PIPELINED && n >= kDiaMinRows)) {
        dia = build_dia(n, npad, rp, cp, vp, index_base, doffs, sv);
        if (!dia && p->matrix_format == 2)
            return CFM_FAIL(CFM_ENOTSUP, "cg: DIA storage needs a banded matrix (<= %d diagonals, sorted rows)",
                            kMaxDiag);
    }
    cfm_status st = CFM_OK;
    if (dia) {
        sp.assign((size_t)npad / 64 + 1, 0);  // unused by the DIA path
        sc.assign(1, -1);
    } else {
        st = build_sell(n, rp, cp, vp, index_base, sp, sc, sv);
        if (st != CFM_OK) return st;
    }

    auto *pl = new (std::nothrow) cfm_cg_plan();
    if (!pl) return CFM_ENOMEM;
    pl->p = *p;
    if (pl->p.max_outer <= 0) pl->p.max_outer = 50;
    if (pl->p.inner_tol <= 0.0) pl->p.inner_tol = 1e-5;  // fastest at 128^3 (2 refinements)
    pl->n = n;
    pl->nnz = nnz;
    pl->ndiag = dia ? (int)doffs.size() : 0;
    pl->nslices = (int)sp.size() - 1;
    pl->nchunks = (n + kChunk - 1) / kChunk;
    auto fail = [&](cfm_status s) {
        cfm_cg_plan_destroy(pl);
        return s;
    };
    st = current_device_props(&pl->dev);
    if (st != CFM_OK) return fail(st);
    // slices are read per chunk: pad the slice table to whole chunks
    const int nslices_pad = pl->nchunks * kSlicesPerChunk;
    const int last_off = sp.back();
    sp.resize(nslices_pad + 1, last_off);
    auto dalloc = [&](void **ptr, size_t bytes) {
        return hipMalloc(ptr, bytes > 0 ? bytes : 16) == hipSuccess;
    };
    const size_t nvec = (size_t)pl->nchunks * kChunk;
    if (!dalloc((void **)&pl->slice_ptr, sizeof(int) * sp.size()) ||
        !dalloc((void **)&pl->scol, sizeof(int) * sc.size()) ||
        !dalloc((void **)&pl->sval, sizeof(double) * sv.size()) ||
        !dalloc((void **)&pl->doff, sizeof(int) * std::max<size_t>(doffs.size(), 1)) ||
        !dalloc((void **)&pl->b, sizeof(double) * nvec) || !dalloc((void **)&pl->x0, sizeof(double) * nvec) ||
        !dalloc((void **)&pl->x, sizeof(double) * nvec) ||
        !dalloc((void **)&pl->part, sizeof(double) * 8 * pl->nchunks) ||
        !dalloc((void **)&pl->scal, sizeof(double) * 2) || !dalloc((void **)&pl->info, sizeof(int) * 4) ||
        !dalloc((void **)&pl->sync, sizeof(GridSync)))
        return fail(CFM_ENOMEM);
    const int nv = p->method == CFM_CG_STANDARD ? 3 : p->method == CFM_CG_PIPELINED ? 6 : 1;
    for (int i = 0; i < nv; ++i)
        if (!dalloc((void **)&pl->v[i], sizeof(double) * nvec)) return fail(CFM_ENOMEM);
    if (p->method == CFM_CG_IR_F32) {
        for (int i = 0; i < 4; ++i)
            if (!dalloc((void **)&pl->f[i], sizeof(float) * nvec)) return fail(CFM_ENOMEM);
        if (!dalloc((void **)&pl->sval32, sizeof(float) * sv.size())) return fail(CFM_ENOMEM);
        std::vector<float> sv32(sv.size());
        for (size_t i = 0; i < sv.size(); ++i) sv32[i] = (float)sv[i];
        if (hipMemcpy(pl->sval32, sv32.data(), sizeof(float) * sv32.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(CFM_EDEVICE);
    }
    if (hipHostMalloc((void **)&pl->scal_h, sizeof(double) * 2) != hipSuccess) return fail(CFM_ENOMEM);
    if (hipMemcpy(pl->slice_ptr, sp.data(), sizeof(int) * sp.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->scol, sc.data(), sizeof(int) * sc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->sval, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (dia && hipMemcpy(pl->doff, doffs.data(), sizeof(int) * doffs.size(), hipMemcpyHostToDevice) != hipSuccess))
        return fail(CFM_EDEVICE);
    if (hipMemset(pl->b, 0, sizeof(double) * nvec) != hipSuccess ||
        hipMemset(pl->x0, 0, sizeof(double) * nvec) != hipSuccess)
        return fail(CFM_EDEVICE);
    if (hipStreamCreate(&pl->stream) != hipSuccess || hipEventCreate(&pl->ev0) != hipSuccess ||
        hipEventCreate(&pl->ev1) != hipSuccess)
        return fail(CFM_EDEVICE);
    const int bt = p->block_threads > 0 ? p->block_threads : 1024;
    if (bt % kThreads != 0 || !kernel_table(bt / kThreads, &pl->K)) return fail(CFM_EINVAL);
    pl->sub = bt / kThreads;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)pl->K.persistent(p->method, p->phase_timers),
                                                     bt, 0) !=
            hipSuccess ||
        per_cu < 1)
        return fail(CFM_EDEVICE);
    const int max_grid = std::min(per_cu * pl->dev.cus, kMaxGrid);
    int want = p->grid_blocks > 0 ? p->grid_blocks : max_grid;
    pl->grid = std::max(1, std::min({want, max_grid, (pl->nchunks + pl->sub - 1) / pl->sub}));
    int rate_khz = 0;
    const unsigned long long us = p->spin_limit_us > 0 ? p->spin_limit_us : 10000000ull;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, pl->dev.device) != hipSuccess ||
        rate_khz <= 0)
        rate_khz = 100000;
    pl->spin_ticks = us * (unsigned long long)rate_khz / 1000ull;
    pl->clock_khz = rate_khz;
    if (p->phase_timers && p->mode == CFM_CG_PERSISTENT &&
        !dalloc((void **)&pl->timers, sizeof(unsigned long long) * kTimerSlots * pl->grid))
        return fail(CFM_ENOMEM);
    *out = pl;
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_plan_set_rhs(cfm_cg_plan *pl, const double *b_host) {
    if (!pl || !b_host) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(pl->b, b_host, sizeof(double) * pl->n, hipMemcpyHostToDevice));
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_plan_set_x0(cfm_cg_plan *pl, const double *x0_host) {
    if (!pl) return CFM_EINVAL;
    if (x0_host)
        CFM_HIP(hipMemcpy(pl->x0, x0_host, sizeof(double) * pl->n, hipMemcpyHostToDevice));
    else
        CFM_HIP(hipMemset(pl->x0, 0, sizeof(double) * pl->n));
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_plan_get_x(cfm_cg_plan *pl, double *x_host) {
    if (!pl || !x_host) return CFM_EINVAL;
    CFM_HIP(hipMemcpy(x_host, pl->x, sizeof(double) * pl->n, hipMemcpyDeviceToHost));
    return CFM_OK;
}

static CgArgs make_args(cfm_cg_plan *pl) {
    CgArgs A{};
    A.n = pl->n;
    A.nchunks = pl->nchunks;
    const int npad = pl->nchunks * kChunk;
    A.A64 = Sell{pl->slice_ptr, pl->scol, pl->sval, pl->doff, pl->ndiag, pl->n, npad};
    A.A32 = Sell{pl->slice_ptr, pl->scol, pl->sval32, pl->doff, pl->ndiag, pl->n, npad};
    A.b = pl->b;
    A.x0 = pl->x0;
    A.x = pl->x;
    for (int i = 0; i < 6; ++i) A.v[i] = pl->v[i];
    for (int i = 0; i < 4; ++i) A.f[i] = pl->f[i];
    A.part = pl->part;
    A.hist = pl->hist;
    A.hist_cap = pl->hist_cap;
    A.max_iter = pl->p.max_iter;
    A.tol = pl->p.tol;
    A.quirk = pl->p.restart_quirk;
    A.max_outer = pl->p.max_outer;
    A.inner_tol = pl->p.inner_tol;
    A.info = pl->info;
    A.sync = pl->sync;
    return A;
}

// Host-driven schedule of the same phases (reference "Baseline Discrete").
// With phase_timers, every operation is bracketed by a roctx range and a
// stream synchronisation and its host time is charged to its phase, as the
// reference's profiling versions bracket theirs with NVTX ranges
// (CG/src/profiling/discrete-standard.cu:222-264).
struct OpTimer {
    bool on;
    hipStream_t s;
    double ms[CFM_NPHASES] = {};
    std::chrono::steady_clock::time_point t0;
    void begin(const char *name) {
        if (!on) return;
        roctxRangePush(name);
        t0 = std::chrono::steady_clock::now();
    }
    cfm_status end(int phase) {
        if (!on) return CFM_OK;
        CFM_HIP(hipStreamSynchronize(s));
        ms[phase] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        roctxRangePop();
        return CFM_OK;
    }
};

static cfm_status run_discrete(cfm_cg_plan *pl, CgArgs &A, std::vector<double> &hist, int *iters_out,
                               int *status_out, OpTimer &ot) {
    const dim3 grid(pl->grid), block(kThreads * pl->sub);
    const CgKernels &K = pl->K;
    hipStream_t s = pl->stream;
    auto fetch = [&](int i0, int i1, double *a, double *b) -> cfm_status {
        ot.begin("Dot (reduce + copy to host)");
        CFM_HIP(launch(K.reduce, dim3(1), block, s, A, i0, i1, pl->scal));
        CFM_HIP(hipMemcpyAsync(pl->scal_h, pl->scal, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
        CFM_HIP(hipStreamSynchronize(s));
        cfm_status st_ = ot.end(CFM_PHASE_REDUCE);
        if (st_ != CFM_OK) return st_;
        *a = pl->scal_h[0];
        if (b) *b = pl->scal_h[1];
        return CFM_OK;
    };
    // launch one operation of the schedule, timed when asked
    auto op = [&](const char *name, int phase, auto launcher) -> cfm_status {
        ot.begin(name);
        CFM_HIP(launcher());
        return ot.end(phase);
    };
    const int max_iter = pl->p.max_iter;
    const double tol = pl->p.tol;
    int k = 0, status = CFM_OK;
    bool converged = false;
    cfm_status st;
    if (pl->p.method == CFM_CG_STANDARD) {
        if ((st = op("Init (r = b - A x0, p = r)", CFM_PHASE_OTHER,
                     [&] { return launch(K.std_init, grid, block, s, A); })) != CFM_OK)
            return st;
        double gamma, bb;
        if ((st = fetch(0, 2, &gamma, &bb)) != CFM_OK) return st;
        const double thr = tol > 0.0 ? tol * tol * bb : 0.0;
        hist.push_back(gamma);
        converged = gamma <= thr;
        while (!converged && k < max_iter) {
            if ((st = op("SpMV (s = A p) + dot <p,s>", CFM_PHASE_SPMV,
                         [&] { return launch(K.std_spmv_dot, grid, block, s, A); })) != CFM_OK)
                return st;
            double delta;
            if ((st = fetch(1, -1, &delta, nullptr)) != CFM_OK) return st;
            if (!(delta > 0.0) || !std::isfinite(delta)) {
                status = CFM_EBREAKDOWN;
                break;
            }
            const double alpha = gamma / delta;
            if ((st = op("Saxpy (x += alpha p, r -= alpha s) + dot <r,r>", CFM_PHASE_UPDATE,
                         [&] { return launch(K.std_update_xr, grid, block, s, A, alpha); })) != CFM_OK)
                return st;
            double gamma_new;
            if ((st = fetch(0, -1, &gamma_new, nullptr)) != CFM_OK) return st;
            ++k;
            hist.push_back(gamma_new);
            if (gamma_new <= thr) {
                converged = true;
                break;
            }
            const double beta = gamma_new / gamma;
            if ((st = op("Scale and saxpy (p = r + beta p)", CFM_PHASE_UPDATE,
                         [&] { return launch(K.std_update_p, grid, block, s, A, beta); })) != CFM_OK)
                return st;
            gamma = gamma_new;
        }
    } else {
        if ((st = op("Init (r = b - A x0)", CFM_PHASE_OTHER,
                     [&] { return launch(K.pipe_init_r, grid, block, s, A); })) != CFM_OK)
            return st;
        if ((st = op("Init (w = A r)", CFM_PHASE_OTHER,
                     [&] { return launch(K.pipe_init_w, grid, block, s, A); })) != CFM_OK)
            return st;
        double bb, dummy;
        if ((st = fetch(4, -1, &bb, nullptr)) != CFM_OK) return st;
        (void)dummy;
        const double thr = tol > 0.0 ? tol * tol * bb : 0.0;
        double alpha = 0.0, beta = 0.0, delta_prev = 0.0;
        for (;;) {
            const int buf = k & 1;
            double delta, gam;
            if ((st = fetch(0 + buf, 2 + buf, &delta, &gam)) != CFM_OK) return st;
            hist.push_back(delta);
            if (delta <= thr) {
                converged = true;
                break;
            }
            if (k >= max_iter) break;
            double denom;
            if (pl->p.restart_quirk ? (k > 1) : (k > 0)) {
                beta = delta / delta_prev;
                denom = gam - (beta / alpha) * delta;
            } else {
                beta = 0.0;
                denom = gam;
            }
            if (!(denom > 0.0) || !std::isfinite(denom)) {
                status = CFM_EBREAKDOWN;
                break;
            }
            alpha = delta / denom;
            if ((st = op("SpMV (q = A w) + dots + vector updates", CFM_PHASE_SPMV,
                         [&] { return launch(K.pipe_iterate, grid, block, s, A, buf, alpha, beta); })) != CFM_OK)
                return st;
            delta_prev = delta;
            ++k;
        }
    }
    CFM_HIP(hipGetLastError());
    if (status == CFM_OK && !converged && tol > 0.0) status = CFM_ENOTCONV;
    *iters_out = k;
    *status_out = status;
    return CFM_OK;
}

// A roctx range around every plan call (a no-op unless a profiler listens).
struct RoctxRange {
    explicit RoctxRange(const char *name) { roctxRangePush(name); }
    ~RoctxRange() { roctxRangePop(); }
};

static const char *solve_range_name(const cfm_cg_params &p) {
    static const char *names[3][2] = {{"cfm_cg_plan_solve standard persistent", "cfm_cg_plan_solve standard discrete"},
                                      {"cfm_cg_plan_solve pipelined persistent", "cfm_cg_plan_solve pipelined discrete"},
                                      {"cfm_cg_plan_solve ir_f32 persistent", "cfm_cg_plan_solve ir_f32 persistent"}};
    return names[p.method][p.mode == CFM_CG_DISCRETE ? 1 : 0];
}

extern "C" cfm_status cfm_cg_plan_phase_times(cfm_cg_plan *pl, cfm_phase_times *out) {
    if (!pl || !out) return CFM_EINVAL;
    if (!pl->p.phase_timers) return CFM_FAIL(CFM_EINVAL, "cg: the plan was created without phase_timers");
    *out = pl->phases;
    return CFM_OK;
}

extern "C" cfm_status cfm_cg_plan_solve(cfm_cg_plan *pl, double *hist, int hist_cap, int *iters,
                                        cfm_timing *t) {
    if (!pl || hist_cap < 0) return CFM_EINVAL;
    RoctxRange range(solve_range_name(pl->p));
    const auto w0 = std::chrono::steady_clock::now();
    // history buffer sized to the request
    const int cap = hist ? hist_cap : 0;
    if (cap > pl->hist_cap) {
        if (pl->hist) (void)hipFree(pl->hist);
        pl->hist = nullptr;
        pl->hist_cap = 0;
        CFM_HIP(hipMalloc((void **)&pl->hist, sizeof(double) * cap));
        pl->hist_cap = cap;
    }
    CgArgs A = make_args(pl);
    A.hist_cap = cap;
    if (!hist) A.hist = nullptr;
    A.timers = pl->timers;
    int status = CFM_OK, k = 0, outer = 0, coop = -1;
    float ms = 0.f;
    const bool timed = pl->p.phase_timers != 0;
    pl->phases = cfm_phase_times{};
    if (pl->p.mode == CFM_CG_PERSISTENT) {
        hipLaunchKernelGGL(reset_sync_kernel, dim3(8), dim3(256), 0, pl->stream, pl->sync, pl->spin_ticks, pl->grid);
        CFM_HIP(hipGetLastError());
        CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
        void *args[] = {&A};
        CFM_HIP(launch_persistent((const void *)pl->K.persistent(pl->p.method, timed), pl->grid, kThreads * pl->sub,
                                  args, pl->stream, &coop));
        CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
        CFM_HIP(hipEventSynchronize(pl->ev1));
        CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
        GridSync hs{};
        CFM_HIP(hipMemcpy(&hs, pl->sync, sizeof(hs), hipMemcpyDeviceToHost));
        if (hs.abort) return barrier_timeout("cg", coop);
        if (timed) {
            std::vector<unsigned long long> ticks((size_t)kTimerSlots * pl->grid);
            CFM_HIP(hipMemcpy(ticks.data(), pl->timers, sizeof(unsigned long long) * ticks.size(),
                              hipMemcpyDeviceToHost));
            summarize_phases(ticks.data(), pl->grid, pl->clock_khz, &pl->phases);
        }
        int info[4] = {0, 0, 0, 0};
        CFM_HIP(hipMemcpy(info, pl->info, sizeof(int) * 4, hipMemcpyDeviceToHost));
        status = info[0];
        k = info[1];
        outer = info[2];
        if (hist && cap > 0) {
            const int nh = pl->p.method == CFM_CG_IR_F32 ? outer + 1 : k + 1;
            CFM_HIP(hipMemcpy(hist, pl->hist, sizeof(double) * std::min(cap, nh), hipMemcpyDeviceToHost));
        }
    } else {
        std::vector<double> h;
        OpTimer ot{timed, pl->stream};
        const auto h0 = std::chrono::steady_clock::now();
        CFM_HIP(hipEventRecord(pl->ev0, pl->stream));
        cfm_status st = run_discrete(pl, A, h, &k, &status, ot);
        if (st != CFM_OK) return st;
        if (timed) {
            for (int i = 0; i < CFM_NPHASES; ++i) pl->phases.ms[i] = pl->phases.max_ms[i] = ot.ms[i];
            pl->phases.total_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            pl->phases.workgroups = 1;
            pl->phases.host_timed = 1;
        }
        CFM_HIP(hipEventRecord(pl->ev1, pl->stream));
        CFM_HIP(hipEventSynchronize(pl->ev1));
        CFM_HIP(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
        if (hist)
            for (int i = 0; i < std::min<int>(cap, (int)h.size()); ++i) hist[i] = h[i];
    }
    if (iters) *iters = k;
    if (t) {
        t->kernel_ms = ms;
        t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        t->grid_blocks = pl->grid;
        t->block_threads = kThreads * pl->sub;
        t->iterations = k;
        t->outer_iterations = outer;
        t->cooperative = coop;
    }
    return (cfm_status)status;
}

extern "C" cfm_status cfm_cg_solve(int n, int nnz, const int *rowptr, const int *col, const double *val,
                                   int index_base, const double *b, double *x_io, const cfm_cg_params *p,
                                   double *hist, int hist_cap, int *iters, cfm_timing *t) {
    if (!b || !x_io) return CFM_EINVAL;
    cfm_cg_plan *pl = nullptr;
    cfm_status st = cfm_cg_plan_create(n, nnz, rowptr, col, val, index_base, 0, p, &pl);
    if (st != CFM_OK) return st;
    st = cfm_cg_plan_set_rhs(pl, b);
    if (st == CFM_OK) st = cfm_cg_plan_set_x0(pl, x_io);
    cfm_status solve_st = CFM_OK;
    if (st == CFM_OK) {
        solve_st = cfm_cg_plan_solve(pl, hist, hist_cap, iters, t);
        if (solve_st == CFM_OK || solve_st == CFM_ENOTCONV || solve_st == CFM_EBREAKDOWN)
            st = cfm_cg_plan_get_x(pl, x_io);
        else
            st = solve_st;
    }
    cfm_cg_plan_destroy(pl);
    return st != CFM_OK ? st : solve_st;
}
