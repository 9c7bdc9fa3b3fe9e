/*
 * CPU oracle: Jacobi 2D 5-point and 3D 7-point sweeps.
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).
 *
 * Restates the reference's single-GPU host loop:
 *   Stencil/jacobi2D/src/common.cu:154-273 (single_gpu)
 *   Stencil/jacobi3D/src/common.cu:110-239 (single_gpu)
 * with the update expressions of
 *   Stencil/jacobi2D/src/common.cu:37-38  new = 0.25 * (E + W + N + S)
 *   Stencil/jacobi3D/src/common.cu:42-46  new = (1/6) * (x+1 + x-1 + y+1 + y-1 + z+1 + z-1)
 * evaluated left to right, no contraction (build with -ffp-contract=off).
 * Periodic mode copies, after every sweep, the last interior row/plane into
 * index 0 and the first interior row/plane into index n-1
 * (Stencil/jacobi2D/src/common.cu:234-242, Stencil/jacobi3D/src/common.cu:198-207),
 * then swaps the arrays; the result is the array named `a` after the final
 * swap (Stencil/jacobi2D/src/common.cu:258).
 *
 * The L2 norm of the update, sqrt(sum (new - old)^2) over the interior every
 * norm_every iterations, follows the convention left commented out at
 * Stencil/jacobi2D/src/common.cu:215-249 (the reference never evaluates it).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Threads of the oracle's OpenMP loops (the CPU baseline times 1 thread and all cores). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

void orc_jacobi_init2d_f64(double *a, int nx, int ny, int kind) {
    /* Stencil/jacobi2D/src/common.cu:15-26 with offset 0: both x faces get
     * sin(2*pi*iy/(ny-1)); everything else 0.  Host glibc sin, pi = 2*asin(1). */
    (void)kind;
    const double pi = 2.0 * asin(1.0);
    memset(a, 0, sizeof(double) * (size_t)nx * ny);
    for (int iy = 0; iy < ny; ++iy) {
        const double y0 = sin(2.0 * pi * iy / (ny - 1));
        a[(size_t)iy * nx + 0] = y0;
        a[(size_t)iy * nx + nx - 1] = y0;
    }
}

void orc_jacobi_init3d_f64(double *a, int nx, int ny, int nz, int kind) {
    const size_t plane = (size_t)nx * ny;
    memset(a, 0, sizeof(double) * plane * nz);
    if (kind == ORC_INIT_REF_QUIRK) {
        /* Stencil/jacobi3D/src/common.cu:18-28: the inner ix loop rewrites the
         * x=0 face; the last write (ix = nx-1) wins: (iz - iy) - (nx-1). */
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                a[iz * plane + (size_t)iy * nx] = ((double)iz - (double)iy) - (double)(nx - 1);
        return;
    }
    const double pi = 2.0 * asin(1.0);
    for (int iz = 0; iz < nz; ++iz) {
        const double sz = sin(2.0 * pi * iz / (nz - 1));
        for (int iy = 0; iy < ny; ++iy) {
            const double v = sin(2.0 * pi * iy / (ny - 1)) * sz;
            a[iz * plane + (size_t)iy * nx + 0] = v;
            a[iz * plane + (size_t)iy * nx + nx - 1] = v;
        }
    }
}

/* FP32 fields: the reference's pi is `real`, i.e. float-rounded for float fields
 * (Stencil/jacobi2D/include/common.h:24, Stencil/jacobi3D/include/common.h:24), and
 * widened back to double in the sin argument; values are then rounded to float. */
void orc_jacobi_init2d_f32(float *a, int nx, int ny, int kind) {
    (void)kind;
    const double pi = (double)(float)(2.0 * asin(1.0));
    memset(a, 0, sizeof(float) * (size_t)nx * ny);
    for (int iy = 0; iy < ny; ++iy) {
        const float y0 = (float)sin(2.0 * pi * iy / (ny - 1));
        a[(size_t)iy * nx + 0] = y0;
        a[(size_t)iy * nx + nx - 1] = y0;
    }
}

void orc_jacobi_init3d_f32(float *a, int nx, int ny, int nz, int kind) {
    const size_t plane = (size_t)nx * ny;
    memset(a, 0, sizeof(float) * plane * nz);
    if (kind == ORC_INIT_REF_QUIRK) {
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                a[iz * plane + (size_t)iy * nx] = (float)(((double)iz - (double)iy) - (double)(nx - 1));
        return;
    }
    const double pi = (double)(float)(2.0 * asin(1.0));
    for (int iz = 0; iz < nz; ++iz) {
        const double sz = sin(2.0 * pi * iz / (nz - 1));
        for (int iy = 0; iy < ny; ++iy) {
            const float v = (float)(sin(2.0 * pi * iy / (ny - 1)) * sz);
            a[iz * plane + (size_t)iy * nx + 0] = v;
            a[iz * plane + (size_t)iy * nx + nx - 1] = v;
        }
    }
}

/* One generic body per element type. */
#define DEFINE_JACOBI2D(NAME, T)                                                              \
    int NAME(const T *init, T *out, int nx, int ny, int niter, int bc, int norm_every,        \
             double *norms, int norm_cap) {                                                   \
        if (nx < 3 || ny < 3 || niter < 0) return -1;                                         \
        const size_t sz = (size_t)nx * ny;                                                    \
        T *a = (T *)malloc(sizeof(T) * sz);                                                   \
        T *an = (T *)malloc(sizeof(T) * sz);                                                  \
        double *rowsum = (double *)malloc(sizeof(double) * ny);                               \
        if (!a || !an || !rowsum) {                                                           \
            free(a);                                                                          \
            free(an);                                                                         \
            free(rowsum);                                                                     \
            return -1;                                                                        \
        }                                                                                     \
        memcpy(a, init, sizeof(T) * sz);                                                      \
        memcpy(an, init, sizeof(T) * sz);                                                     \
        int nnorm = 0;                                                                        \
        for (int iter = 0; iter < niter; ++iter) {                                            \
            const int want = norm_every > 0 && (iter % norm_every) == 0;                      \
            _Pragma("omp parallel for schedule(static)") for (int iy = 1; iy < ny - 1; ++iy) {   \
                double acc = 0.0;                                                             \
                const T *row = a + (size_t)iy * nx;                                           \
                T *dst = an + (size_t)iy * nx;                                                \
                for (int ix = 1; ix < nx - 1; ++ix) {                                         \
                    const T v = (T)0.25 * (row[ix + 1] + row[ix - 1] + row[ix + nx] +         \
                                           row[ix - nx]);                                     \
                    dst[ix] = v;                                                              \
                    if (want) {                                                               \
                        const double d = (double)v - (double)row[ix];                         \
                        acc += d * d;                                                         \
                    }                                                                         \
                }                                                                             \
                rowsum[iy] = acc;                                                             \
            }                                                                                 \
            if (want) {                                                                       \
                double s = 0.0;                                                               \
                for (int iy = 1; iy < ny - 1; ++iy) s += rowsum[iy];                          \
                if (norms && nnorm < norm_cap) norms[nnorm] = sqrt(s);                        \
                nnorm++;                                                                      \
            }                                                                                 \
            if (bc == ORC_BC_PERIODIC) {                                                      \
                memcpy(an, an + (size_t)(ny - 2) * nx, sizeof(T) * nx);                       \
                memcpy(an + (size_t)(ny - 1) * nx, an + (size_t)1 * nx, sizeof(T) * nx);      \
            }                                                                                 \
            T *t = a;                                                                         \
            a = an;                                                                           \
            an = t;                                                                           \
        }                                                                                     \
        memcpy(out, a, sizeof(T) * sz);                                                       \
        free(a);                                                                              \
        free(an);                                                                             \
        free(rowsum);                                                                         \
        return nnorm;                                                                         \
    }

#define DEFINE_JACOBI3D(NAME, T)                                                              \
    int NAME(const T *init, T *out, int nx, int ny, int nz, int niter, int bc, int norm_every, \
             double *norms, int norm_cap) {                                                   \
        if (nx < 3 || ny < 3 || nz < 3 || niter < 0) return -1;                               \
        const size_t plane = (size_t)nx * ny;                                                 \
        const size_t sz = plane * nz;                                                         \
        const T c = (T)1 / (T)6;                                                              \
        T *a = (T *)malloc(sizeof(T) * sz);                                                   \
        T *an = (T *)malloc(sizeof(T) * sz);                                                  \
        double *rowsum = (double *)malloc(sizeof(double) * (size_t)ny * nz);                  \
        if (!a || !an || !rowsum) {                                                           \
            free(a);                                                                          \
            free(an);                                                                         \
            free(rowsum);                                                                     \
            return -1;                                                                        \
        }                                                                                     \
        memcpy(a, init, sizeof(T) * sz);                                                      \
        memcpy(an, init, sizeof(T) * sz);                                                     \
        int nnorm = 0;                                                                        \
        for (int iter = 0; iter < niter; ++iter) {                                            \
            const int want = norm_every > 0 && (iter % norm_every) == 0;                      \
            _Pragma("omp parallel for collapse(2) schedule(static)") for (int iz = 1;          \
                                                                         iz < nz - 1; ++iz) { \
                for (int iy = 1; iy < ny - 1; ++iy) {                                         \
                    double acc = 0.0;                                                         \
                    const T *row = a + iz * plane + (size_t)iy * nx;                          \
                    T *dst = an + iz * plane + (size_t)iy * nx;                               \
                    for (int ix = 1; ix < nx - 1; ++ix) {                                     \
                        const T v = c * (row[ix + 1] + row[ix - 1] + row[ix + nx] +           \
                                         row[ix - nx] + row[ix + plane] + row[ix - plane]);   \
                        dst[ix] = v;                                                          \
                        if (want) {                                                           \
                            const double d = (double)v - (double)row[ix];                     \
                            acc += d * d;                                                     \
                        }                                                                     \
                    }                                                                         \
                    rowsum[(size_t)iz * ny + iy] = acc;                                       \
                }                                                                             \
            }                                                                                 \
            if (want) {                                                                       \
                double s = 0.0;                                                               \
                for (int iz = 1; iz < nz - 1; ++iz)                                           \
                    for (int iy = 1; iy < ny - 1; ++iy) s += rowsum[(size_t)iz * ny + iy];    \
                if (norms && nnorm < norm_cap) norms[nnorm] = sqrt(s);                        \
                nnorm++;                                                                      \
            }                                                                                 \
            if (bc == ORC_BC_PERIODIC) {                                                      \
                memcpy(an, an + (size_t)(nz - 2) * plane, sizeof(T) * plane);                 \
                memcpy(an + (size_t)(nz - 1) * plane, an + plane, sizeof(T) * plane);         \
            }                                                                                 \
            T *t = a;                                                                         \
            a = an;                                                                           \
            an = t;                                                                           \
        }                                                                                     \
        memcpy(out, a, sizeof(T) * sz);                                                       \
        free(a);                                                                              \
        free(an);                                                                             \
        free(rowsum);                                                                         \
        return nnorm;                                                                         \
    }

DEFINE_JACOBI2D(orc_jacobi2d_f64, double)
DEFINE_JACOBI2D(orc_jacobi2d_f32, float)
DEFINE_JACOBI3D(orc_jacobi3d_f64, double)
DEFINE_JACOBI3D(orc_jacobi3d_f32, float)
