/*
 * cpu_baseline.c -- host CG timings for the CPU baseline (TEST INFRASTRUCTURE
 * ONLY: loaded by bench.py's cpu_baseline leg, never by the product library).
 *
 * Two host forms of the same unpreconditioned FP64 CG recurrence as
 * gis/SparseSystem/SolverSingleGPU.h:366-445 (r = b - Ax; p = r; per iteration
 * q = Ap, alpha = rho/(p,q), x += alpha p, r -= alpha q, rho' = (r,r),
 * p = r + (rho'/rho) p), run for a fixed number of iterations:
 *
 *   bl_cg_csr_omp   CSR matvec and BLAS-1 loops parallelised with OpenMP over
 *                   all the threads the caller allows (reduction order left to
 *                   OpenMP: a timing, not a parity, routine);
 *   bl_cg_ldu       one thread, matrix applied in LDU form by the face loop an
 *                   OpenFOAM process uses (y = D x; for each face f:
 *                   y[u_f] += lower_f x[l_f]; y[l_f] += upper_f x[u_f]), as
 *                   the reference's wrappers receive the matrix
 *                   (cuFoamCG/cuFoamCG.C:45-56 hands lduMatrix to the solver).
 *
 * Both return the final ||r||_2 so that the work cannot be optimised away.
 */
#include <math.h>
#include <stdlib.h>
#include <omp.h>

#define BL_API __attribute__((visibility("default")))

BL_API int bl_threads(void) { return omp_get_max_threads(); }

BL_API double bl_cg_csr_omp(int n, const int* rowptr, const int* col, const double* val,
                            const double* b, double* x, int iters) {
    double* r = malloc(sizeof(double) * (size_t)n);
    double* p = malloc(sizeof(double) * (size_t)n);
    double* q = malloc(sizeof(double) * (size_t)n);
    double rho = 0.0;
#pragma omp parallel for reduction(+ : rho) schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s += val[j] * x[col[j]];
        r[i] = b[i] - s;
        p[i] = r[i];
        rho += r[i] * r[i];
    }
    for (int k = 0; k < iters; ++k) {
        double pq = 0.0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s += val[j] * p[col[j]];
            q[i] = s;
            pq += p[i] * s;
        }
        const double alpha = rho / pq;
        double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            rr += r[i] * r[i];
        }
        const double beta = rr / rho;
        rho = rr;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    }
    free(r);
    free(p);
    free(q);
    return sqrt(rho);
}

static void ldu_amul(int n, int m, const int* la, const int* ua, const double* diag,
                     const double* lower, const double* upper, const double* x, double* y) {
    for (int i = 0; i < n; ++i) y[i] = diag[i] * x[i];
    for (int f = 0; f < m; ++f) {
        y[ua[f]] += lower[f] * x[la[f]];
        y[la[f]] += upper[f] * x[ua[f]];
    }
}

BL_API double bl_cg_ldu(int n, int m, const int* la, const int* ua, const double* diag,
                        const double* lower, const double* upper, const double* b, double* x,
                        int iters) {
    double* r = malloc(sizeof(double) * (size_t)n);
    double* p = malloc(sizeof(double) * (size_t)n);
    double* q = malloc(sizeof(double) * (size_t)n);
    ldu_amul(n, m, la, ua, diag, lower, upper, x, q);
    double rho = 0.0;
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        p[i] = r[i];
        rho += r[i] * r[i];
    }
    for (int k = 0; k < iters; ++k) {
        ldu_amul(n, m, la, ua, diag, lower, upper, p, q);
        double pq = 0.0;
        for (int i = 0; i < n; ++i) pq += p[i] * q[i];
        const double alpha = rho / pq;
        double rr = 0.0;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            rr += r[i] * r[i];
        }
        const double beta = rr / rho;
        rho = rr;
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    }
    free(r);
    free(p);
    free(q);
    return sqrt(rho);
}
