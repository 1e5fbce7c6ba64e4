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
 * bl_pcg_openfoam (below) is OpenFOAM's own PCG loop, the CPU solver the
 * wrappers stand in for.
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

/*
 * OpenFOAM's PCG with preconditioner "none", restated from its published
 * algorithm (the CPU solver the gis wrappers replace; OpenFOAM is not part of
 * the cuFoam tree): LDU face-loop matvec, residuals measured as
 * sum|r| / normFactor with normFactor = sum(|A psi - xRef A 1| + |b - xRef A 1|)
 * + 1e-20 (xRef = mean(psi)), stop when the scaled residual falls below
 * `tolerance` (relTol 0) or after max_iter iterations.  One thread.
 * Returns the iteration count; *final_res receives the last scaled residual.
 */
BL_API int bl_pcg_openfoam(int n, int m, const int* la, const int* ua, const double* diag,
                           const double* lower, const double* upper, const double* b,
                           double* x, double tolerance, int max_iter, double* final_res) {
    double* r = malloc(sizeof(double) * (size_t)n);
    double* p = malloc(sizeof(double) * (size_t)n);
    double* w = malloc(sizeof(double) * (size_t)n);
    double* sum_a = malloc(sizeof(double) * (size_t)n);
    ldu_amul(n, m, la, ua, diag, lower, upper, x, w);
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - w[i];
        sum_a[i] = diag[i];   /* A 1, by the same face loop */
    }
    for (int f = 0; f < m; ++f) {
        sum_a[ua[f]] += lower[f];
        sum_a[la[f]] += upper[f];
    }
    double xref = 0.0;
    for (int i = 0; i < n; ++i) xref += x[i];
    xref /= n;
    double nf = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ax = xref * sum_a[i];
        nf += fabs(w[i] - ax) + fabs(b[i] - ax);
    }
    nf += 1e-20;
    double res = 0.0;
    for (int i = 0; i < n; ++i) res += fabs(r[i]);
    res /= nf;
    int it = 0;
    double rho_old = 1.0;
    while (it < max_iter && !(res < tolerance)) {
        double rho = 0.0;
        for (int i = 0; i < n; ++i) rho += r[i] * r[i];
        if (it == 0) {
            for (int i = 0; i < n; ++i) p[i] = r[i];
        } else {
            const double beta = rho / rho_old;
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        }
        rho_old = rho;
        ldu_amul(n, m, la, ua, diag, lower, upper, p, w);
        double pw = 0.0;
        for (int i = 0; i < n; ++i) pw += w[i] * p[i];
        if (fabs(pw) / nf < 1e-300) break;
        const double alpha = rho / pw;
        res = 0.0;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
            res += fabs(r[i]);
        }
        res /= nf;
        ++it;
    }
    *final_res = res;
    free(r);
    free(p);
    free(w);
    free(sum_a);
    return it;
}

/*
 * OpenFOAM's PCG and PBiCG with preconditioner "none" on all host threads.
 * The arithmetic is OpenFOAM's published loops; the matrix is applied row by
 * row from CSR (A and, for PBiCG, A^T) instead of by the serial face loop, as a
 * threaded (or domain-decomposed) run distributes it.  Residuals are
 * sum|r| / normFactor, normFactor = sum(|A psi - xRef A 1| + |b - xRef A 1|)
 * + 1e-20 with xRef = mean(psi); the solve stops when the scaled residual is
 * below `tolerance` (relTol 0) or after max_iter iterations.  Reductions are
 * OpenMP's (a timing routine, not a parity one).  Return the iteration count;
 * *final_res receives the last scaled residual.
 */
static void csr_mul(int n, const int* rowptr, const int* col, const double* val,
                    const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) s += val[j] * x[col[j]];
        y[i] = s;
    }
}

static double omp_norm_factor(int n, const int* rowptr, const double* val, const double* b,
                              const double* x, const double* ax) {
    double xref = 0.0;
#pragma omp parallel for reduction(+ : xref) schedule(static)
    for (int i = 0; i < n; ++i) xref += x[i];
    xref /= n;
    double nf = 0.0;
#pragma omp parallel for reduction(+ : nf) schedule(static)
    for (int i = 0; i < n; ++i) {
        double sum_a = 0.0;   /* (A 1)_i */
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) sum_a += val[j];
        const double xa = xref * sum_a;
        nf += fabs(ax[i] - xa) + fabs(b[i] - xa);
    }
    return nf + 1e-20;
}

BL_API int bl_pcg_openfoam_omp(int n, const int* rowptr, const int* col, const double* val,
                               const double* b, double* x, double tolerance, int max_iter,
                               double* final_res) {
    double* r = malloc(sizeof(double) * (size_t)n);
    double* p = malloc(sizeof(double) * (size_t)n);
    double* w = malloc(sizeof(double) * (size_t)n);
    csr_mul(n, rowptr, col, val, x, w);
    const double nf = omp_norm_factor(n, rowptr, val, b, x, w);
    double res = 0.0;
#pragma omp parallel for reduction(+ : res) schedule(static)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - w[i];
        res += fabs(r[i]);
    }
    res /= nf;
    int it = 0;
    double rho_old = 1.0;
    while (it < max_iter && !(res < tolerance)) {
        double rho = 0.0;
#pragma omp parallel for reduction(+ : rho) schedule(static)
        for (int i = 0; i < n; ++i) rho += r[i] * r[i];
        const double beta = it == 0 ? 0.0 : rho / rho_old;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = it == 0 ? r[i] : r[i] + beta * p[i];
        rho_old = rho;
        csr_mul(n, rowptr, col, val, p, w);
        double pw = 0.0;
#pragma omp parallel for reduction(+ : pw) schedule(static)
        for (int i = 0; i < n; ++i) pw += w[i] * p[i];
        if (fabs(pw) / nf < 1e-300) break;
        const double alpha = rho / pw;
        res = 0.0;
#pragma omp parallel for reduction(+ : res) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
            res += fabs(r[i]);
        }
        res /= nf;
        ++it;
    }
    *final_res = res;
    free(r);
    free(p);
    free(w);
    return it;
}

/* PBiCG, preconditioner none: the shadow system uses A^T (At in CSR) and starts
 * from the same residual, r~ = b - A^T psi as OpenFOAM's Tmul gives it. */
BL_API int bl_pbicg_openfoam_omp(int n, const int* rowptr, const int* col, const double* val,
                                 const int* trowptr, const int* tcol, const double* tval,
                                 const double* b, double* x, double tolerance, int max_iter,
                                 double* final_res) {
    double* rA = malloc(sizeof(double) * (size_t)n);
    double* rT = malloc(sizeof(double) * (size_t)n);
    double* pA = malloc(sizeof(double) * (size_t)n);
    double* pT = malloc(sizeof(double) * (size_t)n);
    double* wA = malloc(sizeof(double) * (size_t)n);
    double* wT = malloc(sizeof(double) * (size_t)n);
    csr_mul(n, rowptr, col, val, x, wA);
    csr_mul(n, trowptr, tcol, tval, x, wT);
    const double nf = omp_norm_factor(n, rowptr, val, b, x, wA);
    double res = 0.0;
#pragma omp parallel for reduction(+ : res) schedule(static)
    for (int i = 0; i < n; ++i) {
        rA[i] = b[i] - wA[i];
        rT[i] = b[i] - wT[i];
        res += fabs(rA[i]);
    }
    res /= nf;
    int it = 0;
    double wArT_old = 0.0;
    while (it < max_iter && !(res < tolerance)) {
        double wArT = 0.0;   /* (r, r~) with the identity preconditioner */
#pragma omp parallel for reduction(+ : wArT) schedule(static)
        for (int i = 0; i < n; ++i) wArT += rA[i] * rT[i];
        const double beta = it == 0 ? 0.0 : wArT / wArT_old;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            pA[i] = it == 0 ? rA[i] : rA[i] + beta * pA[i];
            pT[i] = it == 0 ? rT[i] : rT[i] + beta * pT[i];
        }
        wArT_old = wArT;
        csr_mul(n, rowptr, col, val, pA, wA);
        csr_mul(n, trowptr, tcol, tval, pT, wT);
        double wApT = 0.0;
#pragma omp parallel for reduction(+ : wApT) schedule(static)
        for (int i = 0; i < n; ++i) wApT += wA[i] * pT[i];
        if (fabs(wApT) / nf < 1e-300) break;
        const double alpha = wArT / wApT;
        res = 0.0;
#pragma omp parallel for reduction(+ : res) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * pA[i];
            rA[i] -= alpha * wA[i];
            rT[i] -= alpha * wT[i];
            res += fabs(rA[i]);
        }
        res /= nf;
        ++it;
    }
    *final_res = res;
    free(rA);
    free(rT);
    free(pA);
    free(pT);
    free(wA);
    free(wT);
    return it;
}
