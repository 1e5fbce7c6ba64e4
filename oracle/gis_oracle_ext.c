/*
 * gis_oracle_ext.c -- CPU restatement of the solver paths that have no
 * counterpart in the cuFoam reference (mixed-precision refinement).
 *
 * TEST INFRASTRUCTURE ONLY (see gis_oracle.c for the arithmetic contract:
 * -ffp-contract=off, explicit fma/fmaf, 4-element-leaf pairwise reductions).
 *
 * or_mixed_cg: CG whose Krylov recurrences run in FP32 (CSR values rounded to
 * float, float SpMV products and sequential float row sums, float vector
 * updates with float coefficients) while every reduction accumulates in FP64,
 * and FP64 is used for the residual recompute r = b - A x, the correction
 * accumulation x += (double) xc and the convergence norms.  The stopping rule
 * is the FP64 true relative residual ||b - A x||_2 / ||b||_2 <= rel_tol.
 *
 *   mode 0 (reliable updates): when the FP32 recursive residual estimate falls
 *     below inner_rel_tol times the last true residual (or below the target),
 *     the correction is folded into x in FP64, the true residual is recomputed
 *     in FP64 and replaces the FP32 one, and the search direction continues
 *     with beta = rho_new / rho_old.
 *   mode 1 (restarted refinement): same trigger, but the inner CG restarts
 *     from p = r after every FP64 update.
 *
 * The algorithm is pinned by properties, not by a reference: the returned x
 * satisfies the FP64 true-residual target (recomputed independently by the
 * tests), and mode 0 with inner_rel_tol = 0 reduces to a plain FP32 CG.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))
#define OR_HOT __attribute__((target_clones("fma", "default")))

double or_dot(int64_t n, const double *a, const double *b);
double or_dot_f32(int64_t n, const float *a, const float *b);
void or_residual(int N, const int *rowptr, const int *col, const double *val, const double *b,
                 const double *x, double *r);
void or_spmv_f32(int N, const int *rowptr, const int *col, const float *val, const float *x,
                 float *y);

enum { OR_OK = 0, OR_ERR_NOMEM = 5, OR_NOT_CONVERGED = 11, OR_NAN = 12 };

OR_HOT static void faxpy(int64_t n, float s, const float *w, const float *a, float *v) {
    for (int64_t i = 0; i < n; ++i) v[i] = fmaf(s, w[i], a[i]);
}

OR_API int or_mixed_cg(int N, const int *rowptr, const int *col, const double *val,
                       const double *b, double *x, double rel_tol, double inner_rel_tol,
                       int max_iter, int mode, double *hist, int hist_cap, int *hist_len,
                       int *iters) {
    int64_t nnz = rowptr[N];
    float *val32 = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    double *r = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    float *v32 = (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1) * 4);
    if (!val32 || !r || !v32) { free(val32); free(r); free(v32); return OR_ERR_NOMEM; }
    float *r32 = v32, *p32 = v32 + N, *q32 = v32 + 2 * (size_t)N, *xc = v32 + 3 * (size_t)N;
    for (int64_t j = 0; j < nnz; ++j) val32[j] = (float)val[j];
    int hl = 0;
#define HPUSH(v) do { if (hist && hl < hist_cap) hist[hl] = (v); hl++; } while (0)

    or_residual(N, rowptr, col, val, b, x, r);
    double bnorm = sqrt(or_dot(N, b, b));
    double target = rel_tol * bnorm;
    double rnorm = sqrt(or_dot(N, r, r));
    HPUSH(rnorm);
    int status = OR_OK;
    int k = 0;
    if (rnorm > target) {
        for (int i = 0; i < N; ++i) {
            r32[i] = (float)r[i];
            p32[i] = r32[i];
            xc[i] = 0.0f;
        }
        double rho = or_dot_f32(N, r32, r32);
        double r0 = rnorm;
        int converged = 0;
        while (k < max_iter) {
            ++k;
            or_spmv_f32(N, rowptr, col, val32, p32, q32);
            double pq = or_dot_f32(N, p32, q32);
            double alpha = rho / pq;
            float af = (float)alpha;
            faxpy(N, af, p32, xc, xc);
            faxpy(N, -af, q32, r32, r32);
            double rho_old = rho;
            rho = or_dot_f32(N, r32, r32);
            double est = sqrt(rho);
            if (est != est) { status = OR_NAN; break; }
            if (est < inner_rel_tol * r0 || !(est > target)) {
                for (int i = 0; i < N; ++i) {
                    x[i] = x[i] + (double)xc[i];
                    xc[i] = 0.0f;
                }
                or_residual(N, rowptr, col, val, b, x, r);
                rnorm = sqrt(or_dot(N, r, r));
                HPUSH(rnorm);
                if (!(rnorm > target)) { converged = 1; break; }
                for (int i = 0; i < N; ++i) r32[i] = (float)r[i];
                rho = or_dot_f32(N, r32, r32);
                r0 = rnorm;
                if (mode == 1) {
                    memcpy(p32, r32, sizeof(float) * (size_t)N);
                    continue;
                }
            }
            double beta = rho / rho_old;
            faxpy(N, (float)beta, p32, r32, p32);
        }
        if (!converged) {
            for (int i = 0; i < N; ++i) x[i] = x[i] + (double)xc[i];
            if (status == OR_OK) status = OR_NOT_CONVERGED;
        }
    }
#undef HPUSH
    if (hist_len) *hist_len = hl;
    if (iters) *iters = k;
    free(val32); free(r); free(v32);
    return status;
}

/*
 * or_mixed_krylov: restarted FP32/FP64 refinement of one right-hand side, the
 * per-column arithmetic of gis_solve_multi_mixed (no counterpart in the
 * reference).  Outer step: r64 = b - A x in FP64, ||r64|| (history), stop when
 * ||r64|| <= rel_tol ||b|| (or after max_outer corrections); else solve
 * A d = (float) r64 from d = 0 with an FP32 Krylov method and x += (double) d.
 *
 * The inner methods follow the recurrences and stopping rules of
 * gis/SparseSystem/SolverSingleGPU.h:366-445 (CG, stop when !(||r|| > e)) and
 * :568-675 (BiCG with r* = r, breakdown when |xi| < 1e-10, stop when
 * ||r|| < e), with e = max(inner_rel_tol ||r64||, rel_tol ||b|| / 2), FP32
 * matrix / vectors / updates
 * (float coefficients) and FP64 reductions:
 *   CG:   p = r (k = 1) or fmaf(beta, p, r); x deferred: fmaf(alpha_prev, p_old, x);
 *         q = A32 p; alpha = rho / (p, q); r = fmaf(-alpha, q, r); rho' = (r, r)
 *   BiCG: p = r, p* = r* (k = 1) or fmaf(beta, p, r), fmaf(beta, p*, r*);
 *         q = A32 p; xi = rho / (p*, q); q* = A32^T p*; x = fmaf(xi, p, x);
 *         r = fmaf(-xi, q, r); r* = fmaf(-xi, q*, r*); rho' = (r, r*)
 * Returns the status; hist gets the outer residual norms, *outer their count,
 * *inner the inner iterations summed.
 */
static int inner_krylov(int method, int N, const int *rowptr, const int *col, const float *v32,
                        const int *rowptrT, const int *colT, const float *v32T, const float *b,
                        float *x, double err, int max_inner, float *w) {
    float *r = w, *rs = w + N, *p0 = w + 2 * (size_t)N, *p1 = w + 3 * (size_t)N;
    float *ps = w + 4 * (size_t)N, *q = w + 5 * (size_t)N, *qs = w + 6 * (size_t)N;
    for (int i = 0; i < N; ++i) {
        r[i] = b[i];   /* b - A 0: the products are zeros, the row sums +0 */
        rs[i] = b[i];
        x[i] = 0.0f;
    }
    double rho = or_dot_f32(N, r, r);
    double norm = sqrt(rho);
    if (method == 0 && !(norm > err)) return 0;
    double alpha = 0.0, beta = 0.0;
    int k = 0;
    for (;;) {
        if (k >= max_inner) break;
        const int kk = k + 1;
        float *pold = (kk & 1) ? p0 : p1, *pnew = (kk & 1) ? p1 : p0;
        const float bf = (float)beta, af = (float)alpha;
        for (int i = 0; i < N; ++i) {
            if (kk == 1) {
                pnew[i] = r[i];
                if (method == 1) ps[i] = rs[i];
            } else {
                const float po = pold[i];
                pnew[i] = fmaf(bf, po, r[i]);
                if (method == 1) ps[i] = fmaf(bf, ps[i], rs[i]);
                else x[i] = fmaf(af, po, x[i]);
            }
        }
        or_spmv_f32(N, rowptr, col, v32, pnew, q);
        if (method == 0) {
            alpha = rho / or_dot_f32(N, pnew, q);
            const float a2 = (float)alpha;
            for (int i = 0; i < N; ++i) r[i] = fmaf(-a2, q[i], r[i]);
            const double rr = or_dot_f32(N, r, r);
            beta = rr / rho;
            rho = rr;
            k = kk;
            if (!(sqrt(rr) > err) || k >= max_inner) {
                const float a3 = (float)alpha;   /* the deferred last x update */
                for (int i = 0; i < N; ++i) x[i] = fmaf(a3, pnew[i], x[i]);
                return k;
            }
        } else {
            const double xi = rho / or_dot_f32(N, ps, q);
            if (xi > -1e-10 && xi < 1e-10) return k;   /* breakdown: stop this inner solve */
            or_spmv_f32(N, rowptrT, colT, v32T, ps, qs);
            const float xf = (float)xi;
            for (int i = 0; i < N; ++i) {
                x[i] = fmaf(xf, pnew[i], x[i]);
                r[i] = fmaf(-xf, q[i], r[i]);
                rs[i] = fmaf(-xf, qs[i], rs[i]);
            }
            const double rr = or_dot_f32(N, r, r);
            k = kk;
            if (sqrt(rr) < err || k >= max_inner) return k;
            const double rrs = or_dot_f32(N, r, rs);
            beta = rrs / rho;
            rho = rrs;
        }
    }
    if (method == 0 && k >= 1) {
        const float a3 = (float)alpha;
        float *plast = (k & 1) ? p1 : p0;
        for (int i = 0; i < N; ++i) x[i] = fmaf(a3, plast[i], x[i]);
    }
    return k;
}

OR_API int or_mixed_krylov(int method, int N, const int *rowptr, const int *col,
                           const double *val, const int *rowptrT, const int *colT,
                           const double *valT, const double *b, double *x, double rel_tol,
                           double inner_rel_tol, int max_outer, int max_inner, double *hist,
                           int hist_cap, int *outer, int *inner) {
    int64_t nnz = rowptr[N];
    float *v32 = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1) * 2);
    double *r = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    float *w = (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1) * 9);
    if (!v32 || !r || !w) { free(v32); free(r); free(w); return OR_ERR_NOMEM; }
    float *v32T = v32 + nnz, *r32 = w + 7 * (size_t)N, *d = w + 8 * (size_t)N;
    for (int64_t j = 0; j < nnz; ++j) {
        v32[j] = (float)val[j];
        v32T[j] = valT ? (float)valT[j] : 0.0f;
    }
    const double bnorm = sqrt(or_dot(N, b, b));
    const double target = rel_tol * bnorm;
    int hl = 0, total = 0, status = OR_OK, corrections = 0;
    for (;;) {
        or_residual(N, rowptr, col, val, b, x, r);
        const double rn = sqrt(or_dot(N, r, r));
        if (hist && hl < hist_cap) hist[hl] = rn;
        hl++;
        if (rn != rn) { status = OR_NAN; break; }
        if (!(rn > target)) break;
        if (max_outer > 0 && corrections >= max_outer) { status = OR_NOT_CONVERGED; break; }
        for (int i = 0; i < N; ++i) r32[i] = (float)r[i];
        const double e_in = fmax(inner_rel_tol * rn, 0.5 * target);
        total += inner_krylov(method, N, rowptr, col, v32, rowptrT, colT, v32T, r32, d, e_in,
                              max_inner > 0 ? max_inner : INT_MAX, w);
        for (int i = 0; i < N; ++i) x[i] = x[i] + (double)d[i];
        corrections++;
    }
    *outer = hl;
    *inner = total;
    free(v32);
    free(r);
    free(w);
    return status;
}
