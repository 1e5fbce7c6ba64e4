/*
 * gis_oracle_par.c -- OpenMP restatements of the oracle's CG and mixed CG, for
 * the headline sizes (TEST INFRASTRUCTURE ONLY: tests/, and bench.py's CPU
 * baseline and its error check against an FP64 CPU solve).
 *
 * or_cg_par follows gis/SparseSystem/SolverSingleGPU.h:366-445 exactly as
 * or_cg (gis_oracle.c) does, and or_mixed_cg_par follows or_mixed_cg
 * (gis_oracle_ext.c); both give the same bits as the serial functions, which
 * tests/test_oracle_par.py checks on ragged sizes.  What is parallel:
 *   - CSR row sums: rows are independent, each still summed in storage order;
 *   - vector updates: elementwise;
 *   - reductions: each 1024-row tile (256 leaves, a complete subtree of the
 *     pairwise tree, or_tree.h) is summed on its own thread, then the tile
 *     sums are folded in order by one thread -- the serial tree exactly.
 * Compiled with -ffp-contract=off: fused multiply-adds only where written.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "or_tree.h"

#define OR_API __attribute__((visibility("default")))
#define OR_HOT __attribute__((target_clones("fma", "default")))

enum { OR_OK = 0, OR_ERR_NOMEM = 5, OR_NOT_CONVERGED = 11, OR_NAN = 12 };

#define TILE_ROWS ((int64_t)256 * OR_LEAF)

static inline int64_t ntiles(int64_t n) { return (n + TILE_ROWS - 1) / TILE_ROWS; }

/* ---- per-tile kernels (one call per tile, so the fma clones inline) ---- */
OR_HOT static double tile_dot(int64_t i0, int64_t n, const double *a, const double *b) {
    tree_acc t = {.count = 0};
    const int64_t i1 = i0 + TILE_ROWS < n ? i0 + TILE_ROWS : n;
    for (int64_t i = i0; i < i1; i += OR_LEAF) {
        const int64_t e = i + OR_LEAF < n ? i + OR_LEAF : n;
        double s = 0.0;
        for (int64_t k = i; k < e; ++k) s = fma(a[k], b[k], s);
        tree_push(&t, s);
    }
    return tree_final(&t);
}

OR_HOT static double tile_dot_f32(int64_t i0, int64_t n, const float *a, const float *b) {
    tree_acc t = {.count = 0};
    const int64_t i1 = i0 + TILE_ROWS < n ? i0 + TILE_ROWS : n;
    for (int64_t i = i0; i < i1; i += OR_LEAF) {
        const int64_t e = i + OR_LEAF < n ? i + OR_LEAF : n;
        double s = 0.0;
        for (int64_t k = i; k < e; ++k) s = fma((double)a[k], (double)b[k], s);
        tree_push(&t, s);
    }
    return tree_final(&t);
}

OR_HOT static void tile_axpy(int64_t i0, int64_t i1, double s, const double *w, const double *a,
                             double *v) {
    for (int64_t i = i0; i < i1; ++i) v[i] = fma(s, w[i], a[i]);
}

OR_HOT static void tile_faxpy(int64_t i0, int64_t i1, float s, const float *w, const float *a,
                              float *v) {
    for (int64_t i = i0; i < i1; ++i) v[i] = fmaf(s, w[i], a[i]);
}

static void tile_spmv(int64_t i0, int64_t i1, const int *rowptr, const int *col,
                      const double *val, const double *x, double *y) {
    for (int64_t i = i0; i < i1; ++i) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s = s + val[j] * x[col[j]];
        y[i] = s;
    }
}

static void tile_spmv_f32(int64_t i0, int64_t i1, const int *rowptr, const int *col,
                          const float *val, const float *x, float *y) {
    for (int64_t i = i0; i < i1; ++i) {
        float s = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s = s + val[j] * x[col[j]];
        y[i] = s;
    }
}

static void tile_residual(int64_t i0, int64_t i1, const int *rowptr, const int *col,
                          const double *val, const double *b, const double *x, double *r) {
    for (int64_t i = i0; i < i1; ++i) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s = s + val[j] * x[col[j]];
        r[i] = b[i] - s;
    }
}

/* ---- whole-vector operations ---- */
#define FOR_TILES(n) \
    _Pragma("omp parallel for schedule(static)") \
    for (int64_t t_ = 0; t_ < ntiles(n); ++t_)
#define T0 (t_ * TILE_ROWS)
#define T1(n) (T0 + TILE_ROWS < (n) ? T0 + TILE_ROWS : (n))

static double fold(int64_t nt, const double *part) {
    tree_acc t = {.count = 0};
    for (int64_t i = 0; i < nt; ++i) tree_push(&t, part[i]);
    return tree_final(&t);
}

static double pdot(int64_t n, const double *a, const double *b, double *part) {
    FOR_TILES(n) part[t_] = tile_dot(T0, n, a, b);
    return fold(ntiles(n), part);
}

static double pdot_f32(int64_t n, const float *a, const float *b, double *part) {
    FOR_TILES(n) part[t_] = tile_dot_f32(T0, n, a, b);
    return fold(ntiles(n), part);
}

typedef struct {
    int n;
    const int *rowptr, *col;
} pattern_t;

static void presidual(const pattern_t *A, const double *val, const double *b, const double *x,
                      double *r) {
    FOR_TILES(A->n) tile_residual(T0, T1(A->n), A->rowptr, A->col, val, b, x, r);
}

/* CG (SolverSingleGPU.h:366-445): r = b - A x; p = r; until !(||r|| > error)
 * or max_steps: p = r + (rho/rho_old) p (k > 1); q = A p; xi = rho / (p, q);
 * x += xi p; r -= xi q; rho = (r, r).  History: ||r_k||_2. */
OR_API int or_cg_par(int N, const int *rowptr, const int *col, const double *val,
                     const double *b, double *x, double error, int max_steps, double *hist,
                     int hist_cap, int *hist_len) {
    const pattern_t A = {N, rowptr, col};
    double *r = (double *)malloc(sizeof(double) * ((size_t)N * 3 + (size_t)ntiles(N) + 1));
    if (!r) return OR_ERR_NOMEM;
    double *p = r + N, *q = r + 2 * (size_t)N, *part = r + 3 * (size_t)N;
    int hl = 0;
#define HPUSH(v) do { if (hist && hl < hist_cap) hist[hl] = (v); hl++; } while (0)
    presidual(&A, val, b, x, r);
    FOR_TILES(N) memcpy(p + T0, r + T0, sizeof(double) * (size_t)(T1(N) - T0));
    double rho0 = pdot(N, r, r, part), rho1 = 0.0;
    double norm_r = sqrt(rho0);
    const int64_t ms = max_steps <= 0 ? INT_MAX : max_steps;
    HPUSH(norm_r);
    for (int64_t k = 1; k <= ms && norm_r > error; ++k) {
        if (k > 1) {
            const double beta = rho0 / rho1;
            FOR_TILES(N) tile_axpy(T0, T1(N), beta, p, r, p);
        }
        FOR_TILES(N) tile_spmv(T0, T1(N), rowptr, col, val, p, q);
        const double xi = rho0 / pdot(N, p, q, part);
        FOR_TILES(N) {
            tile_axpy(T0, T1(N), xi, p, x, x);
            tile_axpy(T0, T1(N), -xi, q, r, r);
        }
        rho1 = rho0;
        rho0 = pdot(N, r, r, part);
        norm_r = sqrt(rho0);
        HPUSH(norm_r);
    }
#undef HPUSH
    if (hist_len) *hist_len = hl;
    free(r);
    return OR_OK;
}

/* Mixed-precision CG: or_mixed_cg (gis_oracle_ext.c) step for step. */
OR_API int or_mixed_cg_par(int N, const int *rowptr, const int *col, const double *val,
                           const double *b, double *x, double rel_tol, double inner_rel_tol,
                           int max_iter, int mode, double *hist, int hist_cap, int *hist_len,
                           int *iters) {
    const pattern_t A = {N, rowptr, col};
    const int64_t nnz = rowptr[N];
    const size_t n1 = (size_t)(N > 0 ? N : 1);
    float *val32 = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    double *r = (double *)malloc(sizeof(double) * (n1 + (size_t)ntiles(N) + 1));
    float *v32 = (float *)malloc(sizeof(float) * n1 * 4);
    if (!val32 || !r || !v32) { free(val32); free(r); free(v32); return OR_ERR_NOMEM; }
    double *part = r + n1;
    float *r32 = v32, *p32 = v32 + N, *q32 = v32 + 2 * (size_t)N, *xc = v32 + 3 * (size_t)N;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nnz; ++j) val32[j] = (float)val[j];
    int hl = 0;
#define HPUSH(v) do { if (hist && hl < hist_cap) hist[hl] = (v); hl++; } while (0)
    presidual(&A, val, b, x, r);
    const double bnorm = sqrt(pdot(N, b, b, part));
    const double target = rel_tol * bnorm;
    double rnorm = sqrt(pdot(N, r, r, part));
    HPUSH(rnorm);
    int status = OR_OK;
    int k = 0;
    if (rnorm > target) {
        FOR_TILES(N) {
            for (int64_t i = T0; i < T1(N); ++i) {
                r32[i] = (float)r[i];
                p32[i] = r32[i];
                xc[i] = 0.0f;
            }
        }
        double rho = pdot_f32(N, r32, r32, part);
        double r0 = rnorm;
        int converged = 0;
        while (k < max_iter) {
            ++k;
            FOR_TILES(N) tile_spmv_f32(T0, T1(N), rowptr, col, val32, p32, q32);
            const double pq = pdot_f32(N, p32, q32, part);
            const double alpha = rho / pq;
            const float af = (float)alpha;
            FOR_TILES(N) {
                tile_faxpy(T0, T1(N), af, p32, xc, xc);
                tile_faxpy(T0, T1(N), -af, q32, r32, r32);
            }
            const double rho_old = rho;
            rho = pdot_f32(N, r32, r32, part);
            const double est = sqrt(rho);
            if (est != est) { status = OR_NAN; break; }
            if (est < inner_rel_tol * r0 || !(est > target)) {
                FOR_TILES(N) {
                    for (int64_t i = T0; i < T1(N); ++i) {
                        x[i] = x[i] + (double)xc[i];
                        xc[i] = 0.0f;
                    }
                }
                presidual(&A, val, b, x, r);
                rnorm = sqrt(pdot(N, r, r, part));
                HPUSH(rnorm);
                if (!(rnorm > target)) { converged = 1; break; }
                FOR_TILES(N) {
                    for (int64_t i = T0; i < T1(N); ++i) r32[i] = (float)r[i];
                }
                rho = pdot_f32(N, r32, r32, part);
                r0 = rnorm;
                if (mode == 1) {
                    FOR_TILES(N) memcpy(p32 + T0, r32 + T0, sizeof(float) * (size_t)(T1(N) - T0));
                    continue;
                }
            }
            const float bf = (float)(rho / rho_old);
            FOR_TILES(N) tile_faxpy(T0, T1(N), bf, p32, r32, p32);
        }
        if (!converged) {
            FOR_TILES(N) {
                for (int64_t i = T0; i < T1(N); ++i) x[i] = x[i] + (double)xc[i];
            }
            if (status == OR_OK) status = OR_NOT_CONVERGED;
        }
    }
#undef HPUSH
    if (hist_len) *hist_len = hl;
    if (iters) *iters = k;
    free(val32); free(r); free(v32);
    return status;
}

/* r = b - A x and ||r||_2 (FP64, rows in parallel, the canonical tree). */
OR_API double or_residual_norm_par(int N, const int *rowptr, const int *col, const double *val,
                                   const double *b, const double *x, double *r) {
    const pattern_t A = {N, rowptr, col};
    double *part = (double *)malloc(sizeof(double) * ((size_t)ntiles(N) + 1));
    if (!part) return NAN;
    presidual(&A, val, b, x, r);
    const double nr = sqrt(pdot(N, r, r, part));
    free(part);
    return nr;
}

OR_API int or_par_threads(void);
#ifdef _OPENMP
#include <omp.h>
OR_API int or_par_threads(void) { return omp_get_max_threads(); }
#else
OR_API int or_par_threads(void) { return 1; }
#endif
