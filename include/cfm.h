/*
 * cfm.h -- C ABI of the MI355X-native CPU-free iterative solvers.
 *
 * Plain C types only (pointers, sizes, ints, doubles); no C++ or torch types
 * cross this boundary, nothing here throws, and no entry point calls exit().
 * One GPU per plan.  Plans are not thread-safe; distinct plans are.
 * Host pointers are always owned by the caller; device buffers are owned by
 * the library unless a comment says "borrowed".
 *
 * The reference's operator API is its version table: each version exports
 *     int init(int argc, char **argv)                      (stencils)
 *     int init(int *rowptr_d, int *col_d, real *val_d, int n, int nnz,
 *              bool zero_indexed, int iter_max, real *x_final, ...)  (CG)
 * selected by `-v` (Stencil/jacobi2D/src/main.cu:25-68,
 * Stencil/jacobi3D/src/main.cu:31-72, CG/src/main.cu:24-35,
 * CG/include/common.h:122-133).  Each of those functions parses flags,
 * allocates, launches, times and reports; this ABI splits that work into
 * plan / load / run / download steps so callers can time the kernel alone.
 * The CLI binaries jacobi2D, jacobi3D and cg rebuild the version tables on top.
 */
#ifndef CFM_H
#define CFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CFM_OK = 0,
    CFM_EINVAL = 1,     /* bad argument or unsupported shape */
    CFM_ENOMEM = 2,     /* device or host allocation failed */
    CFM_EDEVICE = 3,    /* no usable GPU, launch or runtime failure */
    CFM_ETIMEOUT = 4,   /* a device-wide barrier exceeded its spin limit */
    CFM_ENOTCONV = 5,   /* CG reached max_iter above the tolerance */
    CFM_EBREAKDOWN = 6, /* CG denominator <= 0 or not finite (matrix not SPD) */
    CFM_ENOTSUP = 7
} cfm_status;

/* Element type of the Jacobi sweep.  CFM_F32 sweeps in FP32 and always
 * accumulates the update norm in FP64. */
enum { CFM_F64 = 0, CFM_F32 = 1 };

/* Handling of the slowest axis (y in 2D, z in 3D).
 * CFM_BC_PERIODIC: after each sweep the last interior row/plane is copied
 *   to index 0 and the first interior row/plane to index n-1, as the
 *   reference's single-GPU loop does (Stencil/jacobi2D/src/common.cu:234-242,
 *   Stencil/jacobi3D/src/common.cu:198-207).
 * CFM_BC_DIRICHLET: every face keeps its initial value. */
enum { CFM_BC_PERIODIC = 0, CFM_BC_DIRICHLET = 1 };

/* Execution scheme of the Jacobi sweep loop. */
enum {
    CFM_JACOBI_LAUNCH_PER_ITER = 0, /* host loop, one launch per sweep ("Baseline Copy") */
    CFM_JACOBI_PERSISTENT = 1,      /* one cooperative launch, device-wide barrier per sweep ("Design 2") */
    CFM_JACOBI_NO_COMPUTE = 2       /* persistent kernel doing only the barriers ("No computation") */
};

/* Built-in initial fields (computed on the host with the C library's sin). */
enum {
    CFM_INIT_SIN = 0,      /* 2D: a[y][0]=a[y][nx-1]=sin(2 pi y/(ny-1)); 3D: both x faces
                              sin(2 pi y/(ny-1)) * sin(2 pi z/(nz-1)); everything else 0 */
    CFM_INIT_REF_QUIRK = 1 /* 3D: x=0 face = (z - y) - (nx-1), the value the reference's
                              P2P build leaves (Stencil/jacobi3D/src/common.cu:15-29) */
};

typedef struct {
    int nx, ny, nz;   /* nz == 1 selects the 2D 5-point stencil; dims >= 3 */
    int niter;        /* sweeps per run, >= 0 */
    int norm_every;   /* > 0: L2 norm of the update on iterations with iter % norm_every == 0 */
    int prec;         /* CFM_F64 or CFM_F32 */
    int bc;           /* CFM_BC_* */
    int variant;      /* CFM_JACOBI_* */
    int tile_rows;    /* rows per wavefront tile (3D: y rows, 2D: rows per prefetch group); 0 = auto */
    int grid_blocks;  /* persistent grid size; 0 = auto (all co-resident workgroups) */
    int chunks;       /* chunks along the streamed axis; 0 = auto */
    int spin_limit_us;/* device barrier watchdog; 0 = default (10 s) */
    int block_threads;/* 0 = one workgroup per CU (size set by the tile shape), or 256 */
    int time_block;   /* sweeps per HBM pass of the persistent kernels: 0 = auto (2D Dirichlet: 4,
                         or 3 with 1024-thread workgroups; 2D periodic y: 2; 3D: 2, Dirichlet or periodic z (periodic:
                         512-thread workgroups; nx > 512: x tiles, which need 16-byte aligned rows, a
                         field below 2 GiB and 512-thread workgroups); with the default tile shape,
                         else 1), 1, 2, or 3 and 4 (2D Dirichlet only; 4 in 512-thread workgroups) */
    int phase_timers; /* nonzero: every workgroup of the persistent kernel records the time it spends
                         sweeping and waiting at the device-wide barrier (cfm_jacobi_plan_phase_times);
                         results are unchanged */
} cfm_jacobi_params;

typedef struct {
    double kernel_ms;   /* HIP-event time of the launches on the plan's stream */
    double wall_ms;     /* host wall time of the run call */
    int grid_blocks;    /* workgroups launched */
    int block_threads;  /* threads per workgroup */
    int iterations;     /* sweeps (Jacobi) or CG iterations (summed inner for IR) */
    int outer_iterations; /* Jacobi: passes over the field in HBM; IR: refinements */
    int cooperative;    /* persistent kernel launch: 1 = cooperative (co-residency guaranteed),
                           0 = ordinary launch because another process held ROCm's one-process
                           cooperative grant (co-residency expected, not guaranteed; the barrier
                           watchdog bounds every wait); -1 = no persistent launch */
} cfm_timing;

/* In-place time breakdown of a run (phase_timers set), per workgroup of the
 * persistent kernel, from its own clock (s_memrealtime) at every phase
 * boundary; the phases of one workgroup add up to its time in the kernel.  No
 * counterpart in the reference, whose breakdown exists only for its
 * host-driven versions (NVTX ranges per operation,
 * CG/src/profiling/discrete-standard.cu:222-264, summed by
 * CG/scripts/measure_operation_breakdown.py). */
enum {
    CFM_PHASE_SPMV = 0,    /* CG: SpMV with its fused dot partials (Pipelined: the whole fused
                              iteration, SpMV + two dots + six updates); Jacobi: the sweeps of a pass */
    CFM_PHASE_UPDATE = 1,  /* CG: vector updates (axpy) with their fused dot partials */
    CFM_PHASE_REDUCE = 2,  /* CG: every workgroup's reduction of the chunk partials to the scalars */
    CFM_PHASE_BARRIER = 3, /* waiting at the device-wide barrier (the slowest workgroup + the barrier) */
    CFM_PHASE_OTHER = 4,   /* set-up: initial residual, refinement outer steps */
    CFM_NPHASES = 5
};
typedef struct {
    double ms[CFM_NPHASES];     /* per phase: mean over workgroups */
    double max_ms[CFM_NPHASES]; /* per phase: the largest workgroup */
    double total_ms;            /* mean over workgroups of the first-to-last clock reading */
    int workgroups;             /* workgroups averaged (0: no breakdown recorded) */
    int host_timed;             /* 1: discrete schedule timed per operation on the host */
} cfm_phase_times;

typedef struct cfm_jacobi_plan cfm_jacobi_plan;
typedef struct cfm_cg_plan cfm_cg_plan;

/* ------------------------------------------------------------ library --- */
const char *cfm_version(void);
const char *cfm_status_str(int status);
/* Message of the last failure recorded on the calling host thread ("" if
 * none): the failed check or HIP call. */
const char *cfm_last_error(void);
cfm_status cfm_device_count(int *count);
/* name, compute units, HBM bytes of device `dev` */
cfm_status cfm_device_info(int dev, char *name, int name_cap, int *compute_units,
                           size_t *mem_bytes);
cfm_status cfm_device_synchronize(void);
/* Streaming copy of `bytes` (two buffers) timed over `reps` launches: the
 * practical HBM ceiling (copy_gbs counts read + write bytes) and the
 * calibration kernel for the FETCH_SIZE / WRITE_SIZE counters.  A utility
 * with no counterpart in the reference. */
cfm_status cfm_bandwidth_probe(size_t bytes, int reps, double *copy_gbs, double *best_ms);
/* One shape of the HBM ceiling sweep over `bytes`: kind 0 copy (gbs counts read + write bytes),
 * 1 read only, 2 write only; `grid` workgroups of 256 lanes, `unroll` (1 or 4) 16-byte vectors
 * in flight per lane, nt nonzero for non-temporal loads and stores.  No counterpart in the
 * reference. */
cfm_status cfm_bandwidth_probe_shape(size_t bytes, int reps, int kind, int grid, int unroll, int nt,
                                     double *gbs, double *best_ms);

/* ------------------------------------------------------------- Jacobi --- */
/* Replaces the allocate/initialise/launch/swap part of
 * SSMultiThreadedTwoBlockComm::init (Stencil/jacobi2D/src/single-stream/
 * multi-threaded-two-block-comm.cu:126-308, Stencil/jacobi3D/src/single-stream/
 * multi-threaded-two-block-comm.cu:126-336) and of single_gpu
 * (Stencil/jacobi2D/src/common.cu:154-273, Stencil/jacobi3D/src/common.cu:110-239). */
cfm_status cfm_jacobi_plan_create(const cfm_jacobi_params *p, cfm_jacobi_plan **out);
/* Copy an nz*ny*nx field of the plan's element type into both working arrays. */
cfm_status cfm_jacobi_plan_load(cfm_jacobi_plan *plan, const void *init_host);
/* Advance the field by niter sweeps.  norms (may be NULL) receives up to
 * norm_cap update norms; norm_count the number produced. */
cfm_status cfm_jacobi_plan_run(cfm_jacobi_plan *plan, double *norms, int norm_cap,
                               int *norm_count, cfm_timing *t);
cfm_status cfm_jacobi_plan_download(cfm_jacobi_plan *plan, void *out_host);
/* Breakdown of the last run of a plan created with phase_timers set. */
cfm_status cfm_jacobi_plan_phase_times(cfm_jacobi_plan *plan, cfm_phase_times *out);
/* Borrowed device pointer to the current field (valid until the next run). */
cfm_status cfm_jacobi_plan_field(cfm_jacobi_plan *plan, void **device_ptr);
void cfm_jacobi_plan_destroy(cfm_jacobi_plan *plan);
/* One shot: create, load, run, download, destroy. */
cfm_status cfm_jacobi_run(const cfm_jacobi_params *p, const void *init_host, void *out_host,
                          double *norms, int norm_cap, int *norm_count, cfm_timing *t);
/* No counterpart in the reference: cfm_jacobi_run for callers that cannot
 * allocate (tight loops, harnesses that forbid hipMalloc).  It makes no
 * hipMalloc / hipFree and creates no stream or event: the barrier state and
 * the update-norm partials live in a 2 MiB device array of the library, one
 * per GPU, held until the process ends.  field (device, borrowed): nz*ny*nx
 * elements of the plan's type, the initial field in, the result out; work
 * (device, borrowed, same size, distinct): overwritten (it starts as a copy
 * of field).  stream: a hipStream_t, NULL for the null stream; the call
 * returns when the run is done.  Same kernels and bits as the plan path.
 * Refused with CFM_EINVAL: phase_timers, and norm slots x (workgroups + 1)
 * beyond the array's 257,888 FP64 words (e.g. 1003 norms at 256 workgroups).
 * Calls that overlap on one GPU share the array and race.  kernel_ms is the
 * host-timed run (no events). */
cfm_status cfm_jacobi_run_noalloc(const cfm_jacobi_params *p, void *field, void *work, void *stream,
                                  double *norms, int norm_cap, int *norm_count, cfm_timing *t);
/* Host generator of the built-in initial fields (prec selects double/float output).
 * Replaces initialize_boundaries (Stencil/jacobi2D/src/common.cu:15-26,
 * Stencil/jacobi3D/src/common.cu:15-29). */
cfm_status cfm_jacobi_init_field(int nx, int ny, int nz, int kind, int prec, void *out_host);

/* ------------------------------------------------------------------ CG --- */
enum { CFM_CG_STANDARD = 0, CFM_CG_PIPELINED = 1, CFM_CG_IR_F32 = 2 };
enum {
    CFM_CG_PERSISTENT = 0, /* whole solve in one cooperative launch ("Single Stream") */
    CFM_CG_DISCRETE = 1    /* one launch per operation + host scalar reads ("Baseline Discrete") */
};

typedef struct {
    int method;          /* CFM_CG_* */
    int mode;            /* CFM_CG_PERSISTENT or CFM_CG_DISCRETE */
    double tol;          /* > 0: stop at <r,r> <= tol^2 <b,b>; <= 0: exactly max_iter iterations */
    int max_iter;        /* CG iterations (IR: inner iterations per refinement) */
    int restart_quirk;   /* pipelined: 1 = reference's `k > 1` restart (default), 0 = `k > 0` */
    int max_outer;       /* IR: refinement steps (0 = 50) */
    double inner_tol;    /* IR: inner relative tolerance (0 = 1e-5), never below half the outer tol */
    int grid_blocks;     /* 0 = auto */
    int spin_limit_us;   /* device barrier watchdog; 0 = default (10 s) */
    int block_threads;   /* 256, 512, 768 or 1024 (256-row sub-blocks per workgroup); 0 = auto */
    int x0_guard;        /* batched plans: nonzero restarts a system from x0 = 0 when
                            ||b - A x0|| > ||b|| (the guard for learned initial guesses) */
    int matrix_format;   /* 0 = auto (single-system Pipelined with >= 2^20 rows: DIA when the CSR
                            is banded with <= 16 diagonals and sorted rows; otherwise SELL-64),
                            1 = SELL-64, 2 = DIA (every system of a batch; ENOTSUP if not banded);
                            an auto batched plan holds SELL-64 and, when every system is banded,
                            DIA beside it for the autotuner to choose (solves start on SELL-64) */
    int phase_timers;    /* nonzero: in-place time breakdown of the solve (cfm_cg_plan_phase_times):
                            the persistent kernel's workgroups record a clock at every phase boundary
                            (a separate kernel build: the default build has no timer code); the
                            discrete schedule times each operation on the host, synchronising after
                            it, and marks it with a roctx range, as the reference's profiling
                            versions do (CG/src/profiling/discrete-standard.cu:222-264).  Results are
                            unchanged (same bits). */
} cfm_cg_params;

/* Replaces SingleStreamStandard::init / SingleStreamPipelined::init and
 * SingleGPUDiscreteStandard::run_single_gpu (CG/include/common.h:117-133,
 * CG/src/single-stream/standard.cu:302-434, CG/src/single-stream/pipelined.cu:321-465,
 * CG/src/common.cu:399-538).  The CSR (int32 indices, float64 values,
 * index_base 0 or 1 as the reference's zero_indexed flag) is read from host
 * memory (csr_on_device = 0) or from borrowed device memory (1) and stored
 * on the device in the library's own sliced layout. */
cfm_status cfm_cg_plan_create(int n, int nnz, const int *rowptr, const int *col,
                              const double *val, int index_base, int csr_on_device,
                              const cfm_cg_params *p, cfm_cg_plan **out);
cfm_status cfm_cg_plan_set_rhs(cfm_cg_plan *plan, const double *b_host);
/* x0 (NULL = zero vector); each solve starts from the stored x0. */
cfm_status cfm_cg_plan_set_x0(cfm_cg_plan *plan, const double *x0_host);
/* Solve from device-resident b and x0.  hist (may be NULL, cap hist_cap)
 * receives <r_k,r_k> for k = 0..iters (IR: the outer residuals). */
cfm_status cfm_cg_plan_solve(cfm_cg_plan *plan, double *hist, int hist_cap, int *iters,
                             cfm_timing *t);
cfm_status cfm_cg_plan_get_x(cfm_cg_plan *plan, double *x_host);
/* Breakdown of the last solve of a plan created with phase_timers set. */
cfm_status cfm_cg_plan_phase_times(cfm_cg_plan *plan, cfm_phase_times *out);
void cfm_cg_plan_destroy(cfm_cg_plan *plan);
/* One shot on host arrays. x_io: x0 in, solution out. */
cfm_status cfm_cg_solve(int n, int nnz, const int *rowptr, const int *col, const double *val,
                        int index_base, const double *b, double *x_io, const cfm_cg_params *p,
                        double *hist, int hist_cap, int *iters, cfm_timing *t);

/* ------------------------------------------------------------ batched --- */
/* Many independent SPD systems in ONE persistent launch (Pipelined CG, one
 * device-wide barrier per iteration, per-system scalars, tolerance and
 * status).  No counterpart in the reference, which solves one matrix per run.
 * The systems form one block-diagonal CSR: rows sys_row_ptr[s] ..
 * sys_row_ptr[s+1]-1 belong to system s and their columns must stay inside
 * that range (0-based).  params: method must be CFM_CG_PIPELINED and mode
 * CFM_CG_PERSISTENT; tol / max_iter / restart_quirk apply to every system.
 * A batch member gives bit for bit the result of solving it alone with
 * cfm_cg_plan_* (same method and flags). */
typedef struct cfm_cg_batch_plan cfm_cg_batch_plan;
cfm_status cfm_cg_batch_plan_create(int nsys, const int64_t *sys_row_ptr, const int *rowptr,
                                    const int *col, const double *val, const cfm_cg_params *p,
                                    cfm_cg_batch_plan **out);
cfm_status cfm_cg_batch_set_rhs(cfm_cg_batch_plan *plan, const double *b);  /* packed, all systems */
cfm_status cfm_cg_batch_set_x0(cfm_cg_batch_plan *plan, const double *x0);  /* NULL = zero */
/* iters/status (may be NULL) receive one entry per system; hist (may be
 * NULL) receives <r_k,r_k> as hist[s * hist_cap + k].  Returns CFM_OK when
 * every system converged, else the worst per-system status. */
cfm_status cfm_cg_batch_solve(cfm_cg_batch_plan *plan, int *iters, int *status, double *hist,
                              int hist_cap, cfm_timing *t);
cfm_status cfm_cg_batch_get_x(cfm_cg_batch_plan *plan, double *x);
/* Autotuner: times trial_iters fixed iterations for each workgroup shape
 * (256/512/1024 threads), grid size (1..occupancy workgroups per CU) and
 * storage the plan holds (SELL-64; DIA too for an auto plan whose systems are
 * all banded), then times full solves (the plan's tol and max_iter, from the
 * current x0) of the three fastest and keeps the fastest solve; cache_path
 * (may be NULL) stores/looks up the winner per (device, system count, row
 * count, tolerance).  The reference has no tuner: its launch shapes are
 * fixed per binary (CG/src/single-stream/pipelined.cu). */
cfm_status cfm_cg_batch_tune(cfm_cg_batch_plan *plan, int trial_iters, const char *cache_path,
                             int *best_block, int *best_grid);
/* Storage the next solve uses: 1 = SELL-64, 2 = DIA (-1 for a NULL plan). */
int cfm_cg_batch_matrix_format(const cfm_cg_batch_plan *plan);
void cfm_cg_batch_plan_destroy(cfm_cg_batch_plan *plan);
/* After a solve with x0_guard set: flags[s] = 1 where system s was restarted
 * from x0 = 0 because its initial guess was worse than zero. */
cfm_status cfm_cg_batch_x0_rejected(cfm_cg_batch_plan *plan, int *flags);

/* ---------------------------------------------------------- generators --- */
/* 3D 7-point Poisson on an nx*ny*nz grid, Dirichlet, rows ix + nx*(iy + ny*iz),
 * columns ascending, diagonal 6, off-diagonals -1 (the CG workload; the
 * reference generates a random tridiagonal instead, CG/src/common.cu:46-76).
 * Returns nnz (>= 0) or -1; pass NULL arrays to query the size. */
int64_t cfm_gen_poisson3d(int nx, int ny, int nz, int *rowptr, int *col, double *val);
/* Variable-coefficient system -div(k grad u) = f, u = 0 outside the box:
 * stream (seed, system_id) draws Lx, Ly, Lz = floor(exp(U(ln lmin, ln(lmax+1))))
 * clipped to [lmin, lmax], then k per cell = exp(N(0,1)) (Box-Muller on
 * uniforms: r = sqrt(-2 ln(1-u1)), z = r cos(2 pi u2), r sin(2 pi u2)), then
 * f per cell = 2u - 1.  7-point rows (ascending columns z-1, y-1, x-1, diag,
 * x+1, y+1, z+1): off-diagonal -2 k_i k_j / (k_i + k_j); a face on the box
 * boundary contributes k_i; the diagonal sums the six face coefficients in
 * that order.  dims receives (Lx, Ly, Lz).  Returns nnz; NULL arrays query
 * the sizes (dims is always written). */
int64_t cfm_gen_varcoef3d(uint64_t seed, int system_id, int lmin, int lmax, int *dims, int *rowptr,
                          int *col, double *val, double *f);
/* The per-cell coefficients k of the same system (kcell holds Lx*Ly*Lz values). */
cfm_status cfm_gen_varcoef3d_coef(uint64_t seed, int system_id, int lmin, int lmax, int *dims,
                                  double *kcell);
/* SplitMix64 stream (seed, stream_id) -> count doubles in [0,1). */
cfm_status cfm_gen_uniform(uint64_t seed, uint64_t stream_id, double *out, int64_t count);

/* ------------------------------------------- the reference's CG inputs --- */
/* The reference's default CG matrix, genTridiag (CG/src/common.cu:46-76; n =
 * 20,971,520 in CG/src/main.cu:118-129): a random symmetric tridiagonal
 * matrix, diagonal rand()/RAND_MAX + 10, off-diagonals rand()/RAND_MAX, drawn
 * in the reference's call order from a private copy of the C library's rand()
 * stream seeded with `seed` (1 = an unseeded program, the reference's case;
 * the caller's rand() state is untouched).  0-based CSR, nnz = 3n - 2.
 * Returns nnz, or -1 for n < 2; NULL arrays query nnz. */
int64_t cfm_gen_tridiag(int n, unsigned int seed, int *rowptr, int *col, double *val);
/* The reference's Matrix Market path: loadMMSparseMatrix
 * (CG/src/mmio_wrapper.cpp:170-369) over mm_read_mtx_crd (CG/src/mmio.c) and
 * main's index-base test (CG/src/main.cu:180-186).  Coordinate files with real
 * or integer data, general / symmetric / skew-symmetric storage (expanded to
 * both triangles); entries sorted by (row, column); pattern, complex, array,
 * repeated entries and both-based indices are refused (CFM_EINVAL with a
 * message), as the reference refuses them; non-square matrices are refused
 * too.  *index_base = 0 if an index 0 occurs, 1 if an index equal to the
 * dimension occurs, else 0; rowptr starts at the base and columns keep it,
 * ready for cfm_cg_plan_create(..., index_base, ...).  The three arrays are
 * allocated by the library; release each with cfm_host_free. */
cfm_status cfm_mm_read(const char *path, int *n, int *nnz, int *index_base, int **rowptr, int **col,
                       double **val);
void cfm_host_free(void *ptr);

/* ---------------------------------------------------- learned warm start --- */
/* No counterpart in the reference: a 3-level 3D U-Net that maps (f, log k)
 * on an n^3 grid to an initial guess u0 for the CG solvers above, trained
 * and run by hand-written BF16 MFMA kernels (FP32 accumulate).
 *
 * Architecture (3x3x3 convolutions, zero padding, ReLU after each):
 *   enc1: 2->16, 16->16        pool 2x2x2 (average)
 *   enc2: 16->32, 32->32       pool
 *   mid : 32->64, 64->64       nearest upsample, concat [up, enc2]
 *   dec2: 96->32, 32->32       nearest upsample, concat [up, enc1]
 *   dec1: 48->16, 16->16       head: 1x1x1 16->1 (FP32, no ReLU)
 * Flat FP32 parameter vector, in layer order: per convolution
 * W[Cout][Cin][27] (tap = 9(dz+1) + 3(dy+1) + (dx+1)) then b[Cout]; then
 * head w[16], head b.  Master parameters stay FP32; forward and backward use
 * BF16 copies of the weights and BF16 activations.
 * input: float [batch][2][n^3] (channel 0 = f, 1 = log k; index (z*n+y)*n+x);
 * output/target: float [batch][n^3].  Loss = mean squared error. */
typedef struct cfm_unet cfm_unet;
size_t cfm_unet_param_count(void);
/* n: multiple of 4, >= 4; batch * n^3 <= 2^27. */
cfm_status cfm_unet_create(int n, int batch, cfm_unet **out);
void cfm_unet_destroy(cfm_unet *net);
/* He-uniform weights (U(-a, a), a = sqrt(6 / (27 Cin))), zero biases, head
 * scaled by 0.1, from a SplitMix64 stream of `seed` (host only). */
cfm_status cfm_unet_init_params(uint64_t seed, float *params);
/* Uploads parameters and resets the Adam state. */
cfm_status cfm_unet_set_params(cfm_unet *net, const float *params);
cfm_status cfm_unet_get_params(cfm_unet *net, float *params);
cfm_status cfm_unet_forward(cfm_unet *net, const float *input, float *output);
/* Forward + backward; grad (may be NULL) receives d loss / d params in the
 * flat order; loss (may be NULL) the mean squared error. */
cfm_status cfm_unet_gradient(cfm_unet *net, const float *input, const float *target, float *grad,
                             double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr. */
cfm_status cfm_unet_train_step(cfm_unet *net, const float *input, const float *target, float lr,
                               double *loss);
/* Network input of every system of a batched plan: float [nsys][2][rows_s],
 * packed in the plan's system order (uploaded once, kept on the device). */
cfm_status cfm_cg_batch_load_features(cfm_cg_batch_plan *plan, const float *features);
/* x0 of every system := scale * net(features), computed and written on the
 * device (no host round trip); every system must be an n^3 cube of the
 * network's n.  Timing: kernel_ms covers the inference and the x0 writes. */
cfm_status cfm_cg_batch_warm_start(cfm_cg_batch_plan *plan, cfm_unet *net, double scale, cfm_timing *t);
/* `sweeps` damped Jacobi sweeps on every system's x0 with the plan's matrix:
 * x <- x + omega (b - A x) / a_ii (0 < omega < 2; rows without a stored
 * diagonal keep x).  Removes the high-frequency error of a learned guess,
 * which inflates ||b - A x0|| (the guard's test) but costs CG little. */
cfm_status cfm_cg_batch_smooth_x0(cfm_cg_batch_plan *plan, int sweeps, double omega, cfm_timing *t);

#ifdef __cplusplus
}
#endif
#endif /* CFM_H */
