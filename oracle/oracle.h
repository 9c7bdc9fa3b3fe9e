/*
 * CPU oracle for the persistent Jacobi and Conjugate Gradient paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libcfm) never links or calls it.
 *
 * Everything here restates the reference's single-GPU host loops in plain C
 * with the exact floating-point operation order that the HIP kernels use:
 *   - Jacobi 2D update      Stencil/jacobi2D/src/common.cu:28-42
 *   - Jacobi 3D update      Stencil/jacobi3D/src/common.cu:31-57
 *   - host loop + periodic  Stencil/jacobi2D/src/common.cu:154-273,
 *     halo copies           Stencil/jacobi3D/src/common.cu:110-239
 *   - Standard CG           CG/src/common.cu:399-538 (run_single_gpu),
 *                           CG/src/single-stream/standard.cu:182-298
 *   - Pipelined CG          CG/src/common.cu:609-780,
 *                           CG/src/single-stream/pipelined.cu:240-317
 *   - CSR SpMV              CG/src/common.cu:150-168
 *
 * Deviations from the reference that the build adopts on purpose:
 *   - an explicit right-hand side b (the reference hard-codes b = 1) and a
 *     relative-residual stopping test (the reference runs a fixed count);
 *   - a deterministic dot-product tree (the reference sums with atomicAdd,
 *     whose order is not reproducible) -- see orc_dot();
 *   - Dirichlet boundary mode beside the reference's periodic halo mode.
 */
#ifndef CFM_ORACLE_H
#define CFM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (numerically identical to the product's cfm_status). */
enum {
    ORC_OK = 0,
    ORC_EINVAL = 1,
    ORC_ENOMEM = 2,
    ORC_ENOTCONV = 5,
    ORC_EBREAKDOWN = 6
};

/* Boundary handling along the slowest axis (y in 2D, z in 3D). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_DIRICHLET = 1 };

/* Initial fields. */
enum {
    ORC_INIT_SIN = 0,      /* 2D: x faces sin(2 pi y/(ny-1)); 3D: x faces sin*sin */
    ORC_INIT_REF_QUIRK = 1 /* 3D only: x=0 face (z - y - (nx-1)), as the reference P2P build */
};

void orc_jacobi_init2d_f64(double *a, int nx, int ny, int kind);
void orc_jacobi_init3d_f64(double *a, int nx, int ny, int nz, int kind);
void orc_jacobi_init2d_f32(float *a, int nx, int ny, int kind);
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_jacobi_init3d_f32(float *a, int nx, int ny, int nz, int kind);

/* Jacobi solvers.  `init` holds the starting field (both working arrays start
 * from it, as the reference initialises a and a_new alike).  `out` receives
 * the field after niter sweeps.  norms (may be NULL) receives
 * sqrt(sum_interior (new-old)^2) for every iteration with iter % norm_every == 0
 * (norm_every <= 0: none); the squared residue is accumulated in double.
 * Returns the number of norms written, or -1 on bad arguments. */
int orc_jacobi2d_f64(const double *init, double *out, int nx, int ny, int niter, int bc,
                     int norm_every, double *norms, int norm_cap);
int orc_jacobi2d_f32(const float *init, float *out, int nx, int ny, int niter, int bc,
                     int norm_every, double *norms, int norm_cap);
int orc_jacobi3d_f64(const double *init, double *out, int nx, int ny, int nz, int niter, int bc,
                     int norm_every, double *norms, int norm_cap);
int orc_jacobi3d_f32(const float *init, float *out, int nx, int ny, int nz, int niter, int bc,
                     int norm_every, double *norms, int norm_cap);

/* ---------------------------------------------------------------- CG ---- */

/* Dot-product tree shared with the GPU kernels: rows are cut into chunks of
 * ORC_DOT_CHUNK; inside a chunk "thread" t (0..255) accumulates rows
 * c*1024 + i*256 + t (i = 0..3) with fma, the 64-lane groups are combined by
 * an xor butterfly (offsets 32,16,8,4,2,1) and the four group sums as
 * (s0+s1)+(s2+s3).  The chunk partials are reduced the same way with
 * thread t adding partials t, t+256, ... in order. */
#define ORC_DOT_CHUNK 1024
#define ORC_DOT_THREADS 256

enum { ORC_TREE_GPU = 0, ORC_TREE_SEQUENTIAL = 1 };

double orc_dot(const double *a, const double *b, int n, int tree);
double orc_dot_f32(const float *a, const float *b, int n, int tree);

/* y = A x, per row sum = fma(val_j, x[col_j], sum) in stored column order.
 * base = 0 or 1 (the reference accepts 1-based Matrix Market CSR). */
void orc_spmv(int n, const int *rowptr, const int *col, const double *val, int base,
              const double *x, double *y);

typedef struct {
    int n;
    const int *rowptr;
    const int *col;
    const double *val;
    int base;
} orc_csr;

/* Standard CG (Hestenes-Stiefel, the reference's schedule).
 * x: x0 on entry, solution on exit.  tol > 0: stop when <r,r> <= tol^2 <b,b>;
 * tol <= 0: run exactly max_iter iterations (the reference's behaviour).
 * hist (cap >= max_iter+1, may be NULL) receives <r_k,r_k> for k = 0..iters.
 * Returns ORC_OK, ORC_ENOTCONV (tol > 0 and max_iter reached) or
 * ORC_EBREAKDOWN (<p,Ap> <= 0 or not finite). */
int orc_cg_standard(const orc_csr *A, const double *b, double *x, double tol, int max_iter,
                    int tree, double *hist, int *iters);

/* Pipelined CG (Ghysels-Vanroose as in the reference).  quirk != 0 keeps the
 * reference's `if (k > 1)` test so that iterations 0 and 1 both restart with
 * beta = 0 (CG/src/single-stream/pipelined.cu:271); quirk == 0 uses k > 0.
 * hist receives <r_k,r_k>. */
int orc_cg_pipelined(const orc_csr *A, const double *b, double *x, double tol, int max_iter,
                     int tree, int quirk, double *hist, int *iters);

/* FP64 iterative refinement around an inner CG whose SpMV and vectors are
 * FP32 (dot products accumulate in FP64).  inner_tol is relative to the
 * outer residual, floored at half the outer target (inner <r,r> <=
 * max(inner_tol^2 <r0,r0>, tol^2 <b,b> / 4)); the outer loop stops when
 * ||b - A x||^2 <= tol^2 ||b||^2.
 * hist receives the outer <r,r> after each refinement (cap max_outer+1);
 * inner_iters_total receives the summed inner iteration count. */
int orc_cg_ir(const orc_csr *A, const double *b, double *x, double tol, int max_outer,
              double inner_tol, int inner_max, double *hist, int *outer_iters,
              int *inner_iters_total);

/* Damped Jacobi smoothing of a guess x in place (sweeps, 0 < omega < 2). */
int orc_jacobi_smooth(const orc_csr *A, const double *b, double *x, int sweeps, double omega);

/* ---------------------------------------------------------- generators -- */

uint64_t orc_splitmix64_next(uint64_t *state);
/* Stream for (seed, stream_id): state0 = seed ^ (0x9E3779B97F4A7C15 * (stream_id + 1)). */
uint64_t orc_stream_state(uint64_t seed, uint64_t stream_id);
/* uniform [0,1): (u >> 11) * 2^-53 */
void orc_uniform(uint64_t seed, uint64_t stream_id, double *out, int64_t count);

/* 3D 7-point Poisson, Dirichlet, rows ix + nx*(iy + ny*iz), columns ascending
 * (z-1, y-1, x-1, diag, x+1, y+1, z+1), diag 6, off-diagonals -1.
 * Returns nnz; rowptr/col/val may be NULL to query the size. */
int64_t orc_poisson3d(int nx, int ny, int nz, int *rowptr, int *col, double *val);

/* Variable-coefficient -div(k grad u) = f system of the batched / warm-start
 * workloads (definition in cg_oracle.c).  Returns nnz; dims receives L. */
int64_t orc_varcoef3d(uint64_t seed, int system_id, int lmin, int lmax, int *dims, int *rowptr, int *col,
                      double *val, double *f);

#ifdef __cplusplus
}
#endif
#endif
