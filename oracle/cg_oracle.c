/*
 * CPU oracle: Conjugate Gradient (Standard, Pipelined, FP32-in-FP64 iterative
 * refinement) and the generators used by the CG configurations.
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).
 *
 * Reference anchors:
 *   SpMV      CG/src/common.cu:150-168   sum += alpha*val*x, alpha = 1, column order
 *   saxpy     CG/src/common.cu:170-176   y = a*x + y
 *   scale+axpy CG/src/common.cu:178-184  y = a*x + scale*y (a = 1)
 *   dot       CG/src/common.cu:357-389   sum of (double)(a_i*b_i)
 *   Standard  CG/src/common.cu:399-538, CG/src/single-stream/standard.cu:182-298
 *   Pipelined CG/src/common.cu:609-780, CG/src/single-stream/pipelined.cu:240-317
 *
 * Contractions: the reference is compiled by nvcc with FMA contraction on,
 * so `a*x + y` and `sum += val*x` become fused multiply-adds.  This oracle
 * (built with -ffp-contract=off) spells every fused operation with fma() and
 * every scalar formula (alpha, beta) without fusion, exactly as the HIP
 * kernels do, so that the two agree bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------ dot tree -- */

static double butterfly64(double *v) {
    /* xor butterfly over 64 lanes; every lane ends with the same value,
     * lane 0 is returned. */
    double t[64];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int i = 0; i < 64; ++i) t[i] = v[i] + v[i ^ off];
        memcpy(v, t, sizeof(t));
    }
    return v[0];
}

static double combine256(double *acc) {
    double s[4];
    for (int w = 0; w < 4; ++w) s[w] = butterfly64(acc + 64 * w);
    return (s[0] + s[1]) + (s[2] + s[3]);
}

static double reduce_partials(const double *part, int nchunks) {
    double acc[ORC_DOT_THREADS];
    for (int t = 0; t < ORC_DOT_THREADS; ++t) {
        double a = 0.0;
        for (int j = t; j < nchunks; j += ORC_DOT_THREADS) a = a + part[j];
        acc[t] = a;
    }
    return combine256(acc);
}

#define DEFINE_DOT(NAME, T)                                                                  \
    double NAME(const T *a, const T *b, int n, int tree) {                                   \
        if (tree == ORC_TREE_SEQUENTIAL) {                                                   \
            double s = 0.0;                                                                  \
            for (int i = 0; i < n; ++i) s = fma((double)a[i], (double)b[i], s);              \
            return s;                                                                        \
        }                                                                                    \
        const int nchunks = (n + ORC_DOT_CHUNK - 1) / ORC_DOT_CHUNK;                         \
        double *part = (double *)malloc(sizeof(double) * (nchunks > 0 ? nchunks : 1));       \
        for (int c = 0; c < nchunks; ++c) {                                                  \
            double acc[ORC_DOT_THREADS];                                                     \
            for (int t = 0; t < ORC_DOT_THREADS; ++t) {                                      \
                double s = 0.0;                                                              \
                for (int i = 0; i < ORC_DOT_CHUNK / ORC_DOT_THREADS; ++i) {                  \
                    const long row = (long)c * ORC_DOT_CHUNK + (long)i * ORC_DOT_THREADS + t; \
                    if (row < n) s = fma((double)a[row], (double)b[row], s);                 \
                }                                                                            \
                acc[t] = s;                                                                  \
            }                                                                                \
            part[c] = combine256(acc);                                                       \
        }                                                                                    \
        const double r = reduce_partials(part, nchunks);                                     \
        free(part);                                                                          \
        return r;                                                                            \
    }

DEFINE_DOT(orc_dot, double)
DEFINE_DOT(orc_dot_f32, float)

/* ---------------------------------------------------------------- SpMV -- */

void orc_spmv(int n, const int *rowptr, const int *col, const double *val, int base,
              const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = rowptr[i] - base; j < rowptr[i + 1] - base; ++j)
            s = fma(val[j], x[col[j] - base], s);
        y[i] = s;
    }
}

static void spmv_f32(int n, const int *rowptr, const int *col, const float *val, int base,
                     const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int j = rowptr[i] - base; j < rowptr[i + 1] - base; ++j)
            s = fmaf(val[j], x[col[j] - base], s);
        y[i] = s;
    }
}

static int bad_csr(const orc_csr *A) {
    return !A || A->n < 1 || !A->rowptr || !A->col || !A->val || (A->base != 0 && A->base != 1);
}

/* ------------------------------------------------------- Standard CG ---- */

int orc_cg_standard(const orc_csr *A, const double *b, double *x, double tol, int max_iter,
                    int tree, double *hist, int *iters) {
    if (bad_csr(A) || !b || !x || max_iter < 0) return ORC_EINVAL;
    const int n = A->n;
    double *r = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *s = (double *)malloc(sizeof(double) * n);
    if (!r || !p || !s) {
        free(r), free(p), free(s);
        return ORC_ENOMEM;
    }
    const double bb = orc_dot(b, b, n, tree);
    const double thr = tol > 0.0 ? tol * tol * bb : 0.0;

    /* r0 = b - A x0 ; p0 = r0  (CG/src/common.cu:442-455) */
    orc_spmv(n, A->rowptr, A->col, A->val, A->base, x, s);
    for (int i = 0; i < n; ++i) {
        r[i] = fma(-1.0, s[i], b[i]);
        p[i] = r[i];
    }
    double gamma = orc_dot(r, r, n, tree);
    if (hist) hist[0] = gamma;

    int k = 0, status = ORC_OK, converged = gamma <= thr;
    while (!converged && k < max_iter) {
        orc_spmv(n, A->rowptr, A->col, A->val, A->base, p, s); /* s = A p */
        const double delta = orc_dot(p, s, n, tree);            /* <p, s>  */
        if (!(delta > 0.0) || !isfinite(delta)) {
            status = ORC_EBREAKDOWN;
            break;
        }
        const double alpha = gamma / delta;
        for (int i = 0; i < n; ++i) {
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, s[i], r[i]);
        }
        const double gamma_new = orc_dot(r, r, n, tree);
        ++k;
        if (hist) hist[k] = gamma_new;
        if (gamma_new <= thr) {
            converged = 1;
            break;
        }
        const double beta = gamma_new / gamma;
        for (int i = 0; i < n; ++i) p[i] = fma(beta, p[i], r[i]);
        gamma = gamma_new;
    }
    if (status == ORC_OK && !converged && tol > 0.0) status = ORC_ENOTCONV;
    if (iters) *iters = k;
    free(r), free(p), free(s);
    return status;
}

/* ------------------------------------------------------ Pipelined CG ---- */

int orc_cg_pipelined(const orc_csr *A, const double *b, double *x, double tol, int max_iter,
                     int tree, int quirk, double *hist, int *iters) {
    if (bad_csr(A) || !b || !x || max_iter < 0) return ORC_EINVAL;
    const int n = A->n;
    double *v[7];
    for (int i = 0; i < 7; ++i) {
        v[i] = (double *)calloc((size_t)n, sizeof(double));
        if (!v[i]) {
            for (int j = 0; j < i; ++j) free(v[j]);
            return ORC_ENOMEM;
        }
    }
    double *r = v[0], *w = v[1], *q = v[2], *z = v[3], *s = v[4], *p = v[5], *ax = v[6];
    const double bb = orc_dot(b, b, n, tree);
    const double thr = tol > 0.0 ? tol * tol * bb : 0.0;

    /* r0 = b - A x0 ; w0 = A r0  (CG/src/common.cu:670-686) */
    orc_spmv(n, A->rowptr, A->col, A->val, A->base, x, ax);
    for (int i = 0; i < n; ++i) r[i] = fma(-1.0, ax[i], b[i]);
    orc_spmv(n, A->rowptr, A->col, A->val, A->base, r, w);
    double delta = orc_dot(r, r, n, tree);
    double gam = orc_dot(r, w, n, tree);
    if (hist) hist[0] = delta;

    double alpha = 0.0, beta = 0.0, delta_prev = 0.0;
    int k = 0, status = ORC_OK, converged = delta <= thr;
    while (!converged && k < max_iter) {
        orc_spmv(n, A->rowptr, A->col, A->val, A->base, w, q); /* q = A w */
        double denom;
        if (quirk ? (k > 1) : (k > 0)) {
            beta = delta / delta_prev;
            denom = gam - (beta / alpha) * delta;
        } else {
            beta = 0.0;
            denom = gam;
        }
        if (!(denom > 0.0) || !isfinite(denom)) {
            status = ORC_EBREAKDOWN;
            break;
        }
        alpha = delta / denom;
        for (int i = 0; i < n; ++i) {
            z[i] = fma(beta, z[i], q[i]);
            s[i] = fma(beta, s[i], w[i]);
            p[i] = fma(beta, p[i], r[i]);
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, s[i], r[i]);
            w[i] = fma(-alpha, z[i], w[i]);
        }
        delta_prev = delta;
        delta = orc_dot(r, r, n, tree);
        gam = orc_dot(r, w, n, tree);
        ++k;
        if (hist) hist[k] = delta;
        if (delta <= thr) converged = 1;
    }
    if (status == ORC_OK && !converged && tol > 0.0) status = ORC_ENOTCONV;
    if (iters) *iters = k;
    for (int i = 0; i < 7; ++i) free(v[i]);
    return status;
}

/* ------------------------------------------ FP32-in-FP64 refinement ---- */

int orc_cg_ir(const orc_csr *A, const double *b, double *x, double tol, int max_outer,
              double inner_tol, int inner_max, double *hist, int *outer_iters,
              int *inner_iters_total) {
    if (bad_csr(A) || !b || !x || max_outer < 0 || inner_max < 1) return ORC_EINVAL;
    const int n = A->n;
    const int nnz = A->rowptr[n] - A->base;
    float *val32 = (float *)malloc(sizeof(float) * (nnz > 0 ? nnz : 1));
    double *r = (double *)malloc(sizeof(double) * n);
    double *ax = (double *)malloc(sizeof(double) * n);
    float *rf = (float *)malloc(sizeof(float) * n);
    float *d = (float *)malloc(sizeof(float) * n);
    float *p = (float *)malloc(sizeof(float) * n);
    float *s = (float *)malloc(sizeof(float) * n);
    if (!val32 || !r || !ax || !rf || !d || !p || !s) {
        free(val32), free(r), free(ax), free(rf), free(d), free(p), free(s);
        return ORC_ENOMEM;
    }
    for (int j = 0; j < nnz; ++j) val32[j] = (float)A->val[j];
    const double bb = orc_dot(b, b, n, ORC_TREE_GPU);
    const double thr = tol * tol * bb;
    int outer = 0, inner_total = 0, status = ORC_ENOTCONV;
    for (;;) {
        /* refinement residual in FP64: r = b - A x */
        orc_spmv(n, A->rowptr, A->col, A->val, A->base, x, ax);
        for (int i = 0; i < n; ++i) r[i] = fma(-1.0, ax[i], b[i]);
        const double rr = orc_dot(r, r, n, ORC_TREE_GPU);
        if (hist) hist[outer] = rr;
        if (rr <= thr) {
            status = ORC_OK;
            break;
        }
        if (outer == max_outer) break;
        /* inner FP32 CG on A d = r, from d = 0 */
        for (int i = 0; i < n; ++i) {
            rf[i] = (float)r[i];
            p[i] = rf[i];
            d[i] = 0.0f;
        }
        double g = orc_dot_f32(rf, rf, n, ORC_TREE_GPU);
        /* no refinement is solved beyond half the outer target */
        const double thr_in = fmax(inner_tol * inner_tol * g, 0.25 * thr);
        int j = 0, brk = 0;
        while (j < inner_max) {
            spmv_f32(n, A->rowptr, A->col, val32, A->base, p, s);
            const double delta = orc_dot_f32(p, s, n, ORC_TREE_GPU);
            if (!(delta > 0.0) || !isfinite(delta)) {
                brk = 1;
                break;
            }
            const float af = (float)(g / delta);
            for (int i = 0; i < n; ++i) {
                d[i] = fmaf(af, p[i], d[i]);
                rf[i] = fmaf(-af, s[i], rf[i]);
            }
            const double gn = orc_dot_f32(rf, rf, n, ORC_TREE_GPU);
            ++j;
            if (gn <= thr_in) break;
            const float bf = (float)(gn / g);
            for (int i = 0; i < n; ++i) p[i] = fmaf(bf, p[i], rf[i]);
            g = gn;
        }
        inner_total += j;
        for (int i = 0; i < n; ++i) x[i] = x[i] + (double)d[i];
        ++outer;
        if (brk && j == 0) {
            status = ORC_EBREAKDOWN;
            break;
        }
    }
    if (outer_iters) *outer_iters = outer;
    if (inner_iters_total) *inner_iters_total = inner_total;
    free(val32), free(r), free(ax), free(rf), free(d), free(p), free(s);
    return status;
}

/* ---------------------------------------------------------- generators -- */

uint64_t orc_splitmix64_next(uint64_t *state) {
    /* SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c). */
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_stream_state(uint64_t seed, uint64_t stream_id) {
    return seed ^ (0x9E3779B97F4A7C15ULL * (stream_id + 1));
}

void orc_uniform(uint64_t seed, uint64_t stream_id, double *out, int64_t count) {
    uint64_t st = orc_stream_state(seed, stream_id);
    for (int64_t i = 0; i < count; ++i)
        out[i] = (double)(orc_splitmix64_next(&st) >> 11) * 0x1.0p-53;
}

int64_t orc_poisson3d(int nx, int ny, int nz, int *rowptr, int *col, double *val) {
    int64_t k = 0;
    const long plane = (long)nx * ny;
    long row = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix, ++row) {
                if (rowptr) rowptr[row] = (int)k;
                const long nb[7] = {iz > 0 ? row - plane : -1,     iy > 0 ? row - nx : -1,
                                    ix > 0 ? row - 1 : -1,         row,
                                    ix < nx - 1 ? row + 1 : -1,    iy < ny - 1 ? row + nx : -1,
                                    iz < nz - 1 ? row + plane : -1};
                for (int j = 0; j < 7; ++j) {
                    if (nb[j] < 0) continue;
                    if (col) col[k] = (int)nb[j];
                    if (val) val[k] = (j == 3) ? 6.0 : -1.0;
                    ++k;
                }
            }
    if (rowptr) rowptr[row] = (int)k;
    return k;
}

/* Variable-coefficient -div(k grad u) = f generator (the batched and learned
 * warm-start workloads).  The definition is the build's own (no reference
 * counterpart): dims ~ floor(exp(U(ln lmin, ln(lmax+1)))) clipped; k =
 * exp(N(0,1)) by Box-Muller; f = 2u - 1; harmonic-mean faces, boundary faces
 * k_i; diagonal = sum of the six face coefficients in z-,y-,x-,x+,y+,z+ order. */
/* separate, non-inlined calls: a compiler may otherwise fuse sin and cos of
 * the same angle into sincos(), whose last bit can differ */
__attribute__((noinline)) static double orc_cos(double a) { return cos(a); }
__attribute__((noinline)) static double orc_sin(double a) { return sin(a); }

int64_t orc_varcoef3d(uint64_t seed, int system_id, int lmin, int lmax, int *dims, int *rowptr, int *col,
                      double *val, double *f) {
    if (!dims || lmin < 1 || lmax < lmin || system_id < 0) return -1;
    uint64_t st = orc_stream_state(seed, (uint64_t)system_id);
    for (int a = 0; a < 3; ++a) {
        const double u = (double)(orc_splitmix64_next(&st) >> 11) * 0x1.0p-53;
        const double lo = log((double)lmin), hi = log((double)lmax + 1.0);
        int v = (int)floor(exp(lo + u * (hi - lo)));
        if (v < lmin) v = lmin;
        if (v > lmax) v = lmax;
        dims[a] = v;
    }
    const long nx = dims[0], ny = dims[1], nz = dims[2], n = nx * ny * nz, plane = nx * ny;
    int64_t nnz = 0;
    for (long iz = 0; iz < nz; ++iz)
        for (long iy = 0; iy < ny; ++iy)
            for (long ix = 0; ix < nx; ++ix)
                nnz += 1 + (iz > 0) + (iy > 0) + (ix > 0) + (ix < nx - 1) + (iy < ny - 1) + (iz < nz - 1);
    if (!rowptr && !col && !val && !f) return nnz;
    double *k = (double *)malloc(sizeof(double) * n);
    if (!k) return -1;
    const double pi = 3.14159265358979323846;
    for (long i = 0; i < n; i += 2) {
        const double u1 = (double)(orc_splitmix64_next(&st) >> 11) * 0x1.0p-53;
        const double u2 = (double)(orc_splitmix64_next(&st) >> 11) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(1.0 - u1));
        k[i] = exp(r * orc_cos(2.0 * pi * u2));
        if (i + 1 < n) k[i + 1] = exp(r * orc_sin(2.0 * pi * u2));
    }
    for (long i = 0; i < n; ++i) {
        const double u = (double)(orc_splitmix64_next(&st) >> 11) * 0x1.0p-53;
        if (f) f[i] = 2.0 * u - 1.0;
    }
    int64_t e = 0;
    long row = 0;
    for (long iz = 0; iz < nz; ++iz)
        for (long iy = 0; iy < ny; ++iy)
            for (long ix = 0; ix < nx; ++ix, ++row) {
                /* neighbours in the fixed face order z-, y-, x-, x+, y+, z+ */
                const long nb[6] = {iz > 0 ? row - plane : -1, iy > 0 ? row - nx : -1, ix > 0 ? row - 1 : -1,
                                    ix < nx - 1 ? row + 1 : -1, iy < ny - 1 ? row + nx : -1,
                                    iz < nz - 1 ? row + plane : -1};
                double c[6], diag = 0.0;
                for (int d = 0; d < 6; ++d) {
                    c[d] = nb[d] >= 0 ? (2.0 * k[row] * k[nb[d]]) / (k[row] + k[nb[d]]) : k[row];
                    diag = diag + c[d];
                }
                if (rowptr) rowptr[row] = (int)e;
                for (int d = 0; d < 3; ++d)
                    if (nb[d] >= 0) {
                        if (col) col[e] = (int)nb[d];
                        if (val) val[e] = -c[d];
                        ++e;
                    }
                if (col) col[e] = (int)row;
                if (val) val[e] = diag;
                ++e;
                for (int d = 3; d < 6; ++d)
                    if (nb[d] >= 0) {
                        if (col) col[e] = (int)nb[d];
                        if (val) val[e] = -c[d];
                        ++e;
                    }
            }
    if (rowptr) rowptr[row] = (int)e;
    free(k);
    return e;
}

/* Damped Jacobi sweeps on a guess (the warm-start smoother of the batched
 * plan, which has no counterpart in the reference).  Row i: ax = SpMV row
 * sum (orc_spmv order), r = fma(-1, ax, b_i), x_i' = fma(omega, r / a_ii, x_i);
 * a row without a stored diagonal keeps x_i. */
int orc_jacobi_smooth(const orc_csr *A, const double *b, double *x, int sweeps, double omega) {
    const int n = A->n, base = A->base;
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!y) return ORC_ENOMEM;
    for (int k = 0; k < sweeps; ++k) {
        for (int i = 0; i < n; ++i) {
            double ax = 0.0, d = 0.0;
            for (int j = A->rowptr[i] - base; j < A->rowptr[i + 1] - base; ++j) {
                const int c = A->col[j] - base;
                ax = fma(A->val[j], x[c], ax);
                if (c == i) d = A->val[j];
            }
            if (d != 0.0) {
                const double r = fma(-1.0, ax, b[i]);
                y[i] = fma(omega, r / d, x[i]);
            } else {
                y[i] = x[i];
            }
        }
        memcpy(x, y, sizeof(double) * (size_t)n);
    }
    free(y);
    return ORC_OK;
}
