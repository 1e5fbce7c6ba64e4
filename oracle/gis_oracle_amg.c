/*
 * gis_oracle_amg.c -- CPU restatement of the preconditioned CG path:
 * preconditioner none / diagonal / GAMG.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product library never does.
 *
 * The reference has no preconditioners (its roadmap, README.md:71-73, lists
 * them as future work); OpenFOAM's PCG with diagonal / GAMG preconditioning
 * is what its users configure instead.  The algorithm below is this
 * project's definition, and the HIP path follows it operation by operation:
 *
 * PCG (gis stopping rule of CG, SolverSingleGPU.h:366-445: ||r||_2 against
 * `error`, history of ||r_k||_2):
 *   r = b - A x; z = M r; rho = (r, z)
 *   k = 1, 2, ...: p = z (k = 1) or fma(beta, p, z); q = A p;
 *                  alpha = rho / (p, q); x = fma(alpha, p, x); r = fma(-alpha, q, r);
 *                  ||r|| = sqrt((r, r)), history, stop test;
 *                  z = M r; rho' = (r, z); beta = rho' / rho.
 *   With M = I this is CG, bit for bit.
 * diagonal:  z_i = (1 / a_ii) r_i.
 * DIC: OpenFOAM's DICPreconditioner, restated from its published algorithm:
 *   rD = diag; for each face f in face order: rD[u] -= upper[f] * upper[f] / rD[l];
 *   rD = 1 / rD;
 *   z = rD .* r; for f in face order: z[u] -= rD[u] * upper[f] * z[l];
 *   for f in reverse face order: z[l] -= rD[l] * upper[f] * z[u].
 * GAMG:
 *   agglomeration of a level: two pairwise passes (pair_pass) over the face
 *     graph with weights |upper_f| + |lower_f|; each pass visits cells in
 *     order and pairs an unpaired cell with its unpaired neighbour of largest
 *     weight (first in row order on ties; rows list the faces where the cell
 *     is neighbour in face order, then those where it is owner), or, with no
 *     unpaired neighbour, joins the aggregate of its heaviest neighbour, or
 *     stays alone.  The second pass runs on the graph of the first pass's
 *     aggregates (weights summed over the merged faces in face order).
 *   coarse level: cells = aggregates; faces = distinct aggregate pairs of
 *     fine faces, in upper-triangular order; Galerkin values (P^T A P with
 *     the 0/1 aggregation P): diag_C = members' diagonals (ascending cell)
 *     then upper + lower of the faces inside C (ascending face);
 *     upper_F / lower_F = the fine face coefficients of F in ascending face
 *     order (swapped where the fine face runs against F).
 *   levels until <= OR_AMG_DENSE cells (dense Gauss-Jordan inverse, no
 *     pivoting), or until a level keeps more than 9/10 of its cells (that
 *     level is then smoothed only: z = sm .* r), at most OR_AMG_LEVELS.
 *   V-cycle at level L with rhs r: sm = omega / diag;
 *     z = sm .* r, then nu - 1 sweeps z = fma(sm, r - A z, z);
 *     t = r - A z; rc_C = sum of t over C's cells (ascending);
 *     z = fma(cscale, zc[agg], z) with zc the V-cycle of level L + 1 on rc;
 *     nu sweeps z = fma(sm, r - A z, z).
 *   Symmetric (same Jacobi sweeps before and after, R = P^T), so PCG applies.
 * GAMG32: the same cycle in FP32 on the same hierarchy: level values
 *   (float) of the FP64 values, sm32 = (float)(omega / diag), the dense
 *   inverse rounded to float, float vectors and float arithmetic (fmaf
 *   where the FP64 cycle has fma, cscale as float); the finest rhs is
 *   (float) r and the result is widened to double; rho = (r, z) in FP64.
 *   The rounding makes the preconditioner vary slightly from one application
 *   to the next, so GAMG32 takes the flexible (Polak-Ribiere) direction
 *   update beta = (rho' - (r', z)) / rho with z the previous preconditioned
 *   residual; (r', z) is 0 in exact arithmetic, where it is CG's beta.
 *
 * Arithmetic contract as in gis_oracle.c (-ffp-contract=off, fma only where
 * written, sequential CSR row sums, the fixed reduction tree of or_dot).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

#define OR_AMG_LEVELS 32
#define OR_AMG_DENSE 64

enum { OR_OK = 0, OR_ERR_ARG = 1, OR_ERR_NOMEM = 5 };
enum { PC_NONE = 0, PC_DIAGONAL = 1, PC_GAMG = 2, PC_GAMG32 = 3, PC_DIC = 4, PC_DILU = 5 };
enum { OR_BREAKDOWN_STATUS = 10 };

/* from gis_oracle.c */
double or_dot(int64_t n, const double *a, const double *b);
int or_ldu2csr_pattern(int N, int M, const int *lower_addr, const int *upper_addr, int *rowptr,
                       int *col, int *perm);
void or_ldu2csr_values(int N, int M, int64_t nnz, const int *perm, const double *diag,
                       const double *lower, const double *upper, double *val);
void or_spmv(int N, const int *rowptr, const int *col, const double *val, const double *x,
             double *y);
void or_spmv_f32(int N, const int *rowptr, const int *col, const float *val, const float *x,
                 float *y);

typedef struct {
    int n, m;
    int *la, *ua;
    double *diag, *lower, *upper;
    int *rowptr, *col, *perm;
    double *val;
    double *sm;
    int nc;      /* cells of the next level (0 on the last level) */
    int *agg;    /* cell -> next-level cell */
    int *fcode;  /* face -> 2 F + flip (F a next-level face) or -1 - C (inside next-level cell C) */
    double *r, *z, *z2, *t;  /* V-cycle work vectors */
    float *val32, *sm32, *r32, *z32, *z232, *t32;   /* the FP32 cycle (GAMG32) */
} level_t;

typedef struct {
    int nl;
    level_t lv[OR_AMG_LEVELS];
    double *ainv;  /* dense inverse of the last level, or NULL */
    float *ainv32; /* (float) ainv, made with the FP32 cycle's data */
    double omega, cscale;
    int nu;
} amg_t;

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

/* rows of the face graph: faces where the cell is neighbour (face order),
 * then faces where it is owner (face order); other cell in nb */
static int graph_rows(int n, int m, const int *la, const int *ua, int **rp_out, int **face_out,
                      int **nb_out) {
    int *rp = xcalloc((size_t)n + 1, sizeof(int));
    int *face = xcalloc(2 * (size_t)m, sizeof(int));
    int *nb = xcalloc(2 * (size_t)m, sizeof(int));
    int *lo_cnt = xcalloc((size_t)n, sizeof(int));
    int *fill = xcalloc((size_t)n, sizeof(int));
    if (!rp || !face || !nb || !lo_cnt || !fill) {
        free(rp); free(face); free(nb); free(lo_cnt); free(fill);
        return OR_ERR_NOMEM;
    }
    for (int f = 0; f < m; ++f) {
        rp[ua[f] + 1]++;
        rp[la[f] + 1]++;
        lo_cnt[ua[f]]++;
    }
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    for (int f = 0; f < m; ++f) {   /* neighbour part */
        const int c = ua[f];
        const int p = rp[c] + fill[c]++;
        face[p] = f;
        nb[p] = la[f];
    }
    for (int i = 0; i < n; ++i) fill[i] = 0;
    for (int f = 0; f < m; ++f) {   /* owner part */
        const int c = la[f];
        const int p = rp[c] + lo_cnt[c] + fill[c]++;
        face[p] = f;
        nb[p] = ua[f];
    }
    free(lo_cnt);
    free(fill);
    *rp_out = rp;
    *face_out = face;
    *nb_out = nb;
    return OR_OK;
}

/* one pairwise pass; returns the aggregate count or -1 (no memory) */
static int pair_pass(int n, int m, const int *la, const int *ua, const double *w, int *agg) {
    int *rp, *face, *nb;
    if (graph_rows(n, m, la, ua, &rp, &face, &nb) != OR_OK) return -1;
    for (int i = 0; i < n; ++i) agg[i] = -1;
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int best = -1;
        double bw = -1.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (agg[nb[p]] < 0 && w[face[p]] > bw) {
                bw = w[face[p]];
                best = nb[p];
            }
        if (best >= 0) {
            agg[i] = agg[best] = nc++;
            continue;
        }
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (w[face[p]] > bw) {
                bw = w[face[p]];
                best = nb[p];
            }
        agg[i] = best >= 0 ? agg[best] : nc++;
    }
    free(rp);
    free(face);
    free(nb);
    return nc;
}

/* one pairwise pass, exported for known-answer tests */
OR_API int or_pair_pass(int n, int m, const int *la, const int *ua, const double *w, int *agg) {
    return pair_pass(n, m, la, ua, w, agg);
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* faces of the aggregate graph; fcode as in level_t; returns the face count
 * (or -1) and the addressing in *cla, *cua */
static int coarse_faces(int m, const int *la, const int *ua, const int *agg, int nc,
                        int *fcode, int **cla, int **cua) {
    int64_t *keys = xcalloc((size_t)m, sizeof(int64_t));
    if (!keys) return -1;
    int nk = 0;
    for (int f = 0; f < m; ++f) {
        const int a = agg[la[f]], b = agg[ua[f]];
        if (a != b) keys[nk++] = (int64_t)(a < b ? a : b) * nc + (a < b ? b : a);
    }
    qsort(keys, (size_t)nk, sizeof(int64_t), cmp_i64);
    int mc = 0;
    for (int k = 0; k < nk; ++k)
        if (k == 0 || keys[k] != keys[k - 1]) keys[mc++] = keys[k];
    *cla = xcalloc((size_t)mc, sizeof(int));
    *cua = xcalloc((size_t)mc, sizeof(int));
    if (!*cla || !*cua) {
        free(keys);
        return -1;
    }
    for (int F = 0; F < mc; ++F) {
        (*cla)[F] = (int)(keys[F] / nc);
        (*cua)[F] = (int)(keys[F] % nc);
    }
    for (int f = 0; f < m; ++f) {
        const int a = agg[la[f]], b = agg[ua[f]];
        if (a == b) {
            fcode[f] = -1 - a;
            continue;
        }
        const int64_t key = (int64_t)(a < b ? a : b) * nc + (a < b ? b : a);
        int lo = 0, hi = mc - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1; else hi = mid;
        }
        fcode[f] = 2 * lo + (a > b ? 1 : 0);
    }
    free(keys);
    return mc;
}

static int level_csr(level_t *L) {
    const int64_t nnz = (int64_t)L->n + 2 * (int64_t)L->m;
    L->rowptr = xcalloc((size_t)L->n + 1, sizeof(int));
    L->col = xcalloc((size_t)nnz, sizeof(int));
    L->perm = xcalloc((size_t)nnz, sizeof(int));
    L->val = xcalloc((size_t)nnz, sizeof(double));
    if (!L->rowptr || !L->col || !L->perm || !L->val) return OR_ERR_NOMEM;
    return or_ldu2csr_pattern(L->n, L->m, L->la, L->ua, L->rowptr, L->col, L->perm);
}

static void level_values(level_t *L, double omega) {
    const int64_t nnz = (int64_t)L->n + 2 * (int64_t)L->m;
    or_ldu2csr_values(L->n, L->m, nnz, L->perm, L->diag, L->lower, L->upper, L->val);
    for (int i = 0; i < L->n; ++i) L->sm[i] = omega / L->diag[i];
}

/* Galerkin values of level L+1 from level L */
static void restrict_values(const level_t *F, level_t *C) {
    for (int c = 0; c < C->n; ++c) C->diag[c] = 0.0;
    for (int g = 0; g < C->m; ++g) C->upper[g] = C->lower[g] = 0.0;
    for (int i = 0; i < F->n; ++i) C->diag[F->agg[i]] = C->diag[F->agg[i]] + F->diag[i];
    for (int f = 0; f < F->m; ++f) {
        const int code = F->fcode[f];
        if (code < 0) {
            const int c = -1 - code;
            C->diag[c] = C->diag[c] + F->upper[f];
            C->diag[c] = C->diag[c] + F->lower[f];
        } else {
            const int g = code >> 1;
            const int flip = code & 1;
            C->upper[g] = C->upper[g] + (flip ? F->lower[f] : F->upper[f]);
            C->lower[g] = C->lower[g] + (flip ? F->upper[f] : F->lower[f]);
        }
    }
}

static void free_level(level_t *L) {
    free(L->la); free(L->ua); free(L->diag); free(L->lower); free(L->upper);
    free(L->rowptr); free(L->col); free(L->perm); free(L->val); free(L->sm);
    free(L->agg); free(L->fcode); free(L->r); free(L->z); free(L->z2); free(L->t);
    free(L->val32); free(L->sm32); free(L->r32); free(L->z32); free(L->z232); free(L->t32);
}

static int alloc_level(level_t *L, int n, int m) {
    memset(L, 0, sizeof(*L));
    L->n = n;
    L->m = m;
    L->la = xcalloc((size_t)m, sizeof(int));
    L->ua = xcalloc((size_t)m, sizeof(int));
    L->diag = xcalloc((size_t)n, sizeof(double));
    L->lower = xcalloc((size_t)m, sizeof(double));
    L->upper = xcalloc((size_t)m, sizeof(double));
    L->sm = xcalloc((size_t)n, sizeof(double));
    L->r = xcalloc((size_t)n, sizeof(double));
    L->z = xcalloc((size_t)n, sizeof(double));
    L->z2 = xcalloc((size_t)n, sizeof(double));
    L->t = xcalloc((size_t)n, sizeof(double));
    if (!L->la || !L->ua || !L->diag || !L->lower || !L->upper || !L->sm || !L->r || !L->z ||
        !L->z2 || !L->t)
        return OR_ERR_NOMEM;
    return OR_OK;
}

/* aggregation of level L (two pairwise passes); fills L->agg, L->fcode, L->nc
 * and returns the next level's addressing */
static int aggregate(level_t *L, int **cla, int **cua, int *mc) {
    const int n = L->n, m = L->m;
    double *w = xcalloc((size_t)m, sizeof(double));
    int *agg1 = xcalloc((size_t)n, sizeof(int));
    int *fc1 = xcalloc((size_t)m, sizeof(int));
    L->agg = xcalloc((size_t)n, sizeof(int));
    L->fcode = xcalloc((size_t)m, sizeof(int));
    if (!w || !agg1 || !fc1 || !L->agg || !L->fcode) return OR_ERR_NOMEM;
    for (int f = 0; f < m; ++f) w[f] = fabs(L->upper[f]) + fabs(L->lower[f]);
    const int n1 = pair_pass(n, m, L->la, L->ua, w, agg1);
    int *la1, *ua1;
    const int m1 = coarse_faces(m, L->la, L->ua, agg1, n1, fc1, &la1, &ua1);
    if (n1 < 0 || m1 < 0) return OR_ERR_NOMEM;
    double *w1 = xcalloc((size_t)m1, sizeof(double));
    int *agg2 = xcalloc((size_t)n1, sizeof(int));
    if (!w1 || !agg2) return OR_ERR_NOMEM;
    for (int f = 0; f < m; ++f)
        if (fc1[f] >= 0) w1[fc1[f] >> 1] = w1[fc1[f] >> 1] + w[f];
    const int n2 = pair_pass(n1, m1, la1, ua1, w1, agg2);
    if (n2 < 0) return OR_ERR_NOMEM;
    for (int i = 0; i < n; ++i) L->agg[i] = agg2[agg1[i]];
    L->nc = n2;
    *mc = coarse_faces(m, L->la, L->ua, L->agg, n2, L->fcode, cla, cua);
    free(w); free(agg1); free(fc1); free(la1); free(ua1); free(w1); free(agg2);
    return *mc < 0 ? OR_ERR_NOMEM : OR_OK;
}

static void dense_inverse(const level_t *L, double *ainv) {
    const int n = L->n, w = 2 * n;
    double *M = xcalloc((size_t)n * w, sizeof(double));
    for (int i = 0; i < n; ++i) {
        M[(size_t)i * w + i] = L->diag[i];
        M[(size_t)i * w + n + i] = 1.0;
    }
    for (int f = 0; f < L->m; ++f) {
        M[(size_t)L->la[f] * w + L->ua[f]] = M[(size_t)L->la[f] * w + L->ua[f]] + L->upper[f];
        M[(size_t)L->ua[f] * w + L->la[f]] = M[(size_t)L->ua[f] * w + L->la[f]] + L->lower[f];
    }
    double *fac = xcalloc((size_t)n, sizeof(double));
    for (int k = 0; k < n; ++k) {
        const double piv = M[(size_t)k * w + k];
        for (int j = 0; j < w; ++j) M[(size_t)k * w + j] = M[(size_t)k * w + j] / piv;
        for (int i = 0; i < n; ++i) fac[i] = M[(size_t)i * w + k];
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            for (int j = 0; j < w; ++j)
                M[(size_t)i * w + j] = fma(-fac[i], M[(size_t)k * w + j], M[(size_t)i * w + j]);
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) ainv[(size_t)i * n + j] = M[(size_t)i * w + n + j];
    free(M);
    free(fac);
}

OR_API void or_amg_free(void *p) {
    amg_t *a = (amg_t *)p;
    if (!a) return;
    for (int l = 0; l < a->nl; ++l) free_level(&a->lv[l]);
    free(a->ainv);
    free(a->ainv32);
    free(a);
}

/* Hierarchy of an LDU system; NULL on failure. */
OR_API void *or_amg_build(int N, int M, const int *la, const int *ua, const double *diag,
                          const double *lower, const double *upper, double omega, int nu,
                          double cscale) {
    if (N < 1 || M < 0 || nu < 1) return NULL;
    amg_t *a = xcalloc(1, sizeof(amg_t));
    if (!a) return NULL;
    a->omega = omega;
    a->nu = nu;
    a->cscale = cscale;
    level_t *L = &a->lv[0];
    a->nl = 1;
    if (alloc_level(L, N, M) != OR_OK) goto fail;
    memcpy(L->la, la, sizeof(int) * (size_t)M);
    memcpy(L->ua, ua, sizeof(int) * (size_t)M);
    memcpy(L->diag, diag, sizeof(double) * (size_t)N);
    memcpy(L->lower, lower, sizeof(double) * (size_t)M);
    memcpy(L->upper, upper, sizeof(double) * (size_t)M);
    if (level_csr(L) != OR_OK) goto fail;
    level_values(L, omega);
    while (L->n > OR_AMG_DENSE && a->nl < OR_AMG_LEVELS) {
        int *cla = NULL, *cua = NULL, mc = 0;
        if (aggregate(L, &cla, &cua, &mc) != OR_OK) goto fail;
        if ((int64_t)L->nc * 10 > (int64_t)L->n * 9) {   /* stagnation: stop here */
            free(cla); free(cua);
            free(L->agg); free(L->fcode);
            L->agg = L->fcode = NULL;
            L->nc = 0;
            break;
        }
        level_t *C = &a->lv[a->nl];
        if (alloc_level(C, L->nc, mc) != OR_OK) goto fail;
        a->nl++;
        memcpy(C->la, cla, sizeof(int) * (size_t)mc);
        memcpy(C->ua, cua, sizeof(int) * (size_t)mc);
        free(cla);
        free(cua);
        restrict_values(L, C);
        if (level_csr(C) != OR_OK) goto fail;
        level_values(C, omega);
        L = C;
    }
    if (L->n <= OR_AMG_DENSE) {
        a->ainv = xcalloc((size_t)L->n * L->n, sizeof(double));
        if (!a->ainv) goto fail;
        dense_inverse(L, a->ainv);
    }
    return a;
fail:
    or_amg_free(a);
    return NULL;
}

OR_API int or_amg_levels(void *p) { return p ? ((amg_t *)p)->nl : 0; }

/* sizes of level l: n, m, nc (next-level cells, 0 on the last) */
OR_API int or_amg_level_size(void *p, int l, int *n, int *m, int *nc) {
    amg_t *a = (amg_t *)p;
    if (!a || l < 0 || l >= a->nl) return OR_ERR_ARG;
    *n = a->lv[l].n;
    *m = a->lv[l].m;
    *nc = a->lv[l].nc;
    return OR_OK;
}

/* copies of level l (any pointer may be NULL; agg / fcode only below the last level) */
OR_API int or_amg_level_get(void *p, int l, int *la, int *ua, double *diag, double *lower,
                            double *upper, int *agg, int *fcode) {
    amg_t *a = (amg_t *)p;
    if (!a || l < 0 || l >= a->nl) return OR_ERR_ARG;
    const level_t *L = &a->lv[l];
    if (la) memcpy(la, L->la, sizeof(int) * (size_t)L->m);
    if (ua) memcpy(ua, L->ua, sizeof(int) * (size_t)L->m);
    if (diag) memcpy(diag, L->diag, sizeof(double) * (size_t)L->n);
    if (lower) memcpy(lower, L->lower, sizeof(double) * (size_t)L->m);
    if (upper) memcpy(upper, L->upper, sizeof(double) * (size_t)L->m);
    if (agg && L->agg) memcpy(agg, L->agg, sizeof(int) * (size_t)L->n);
    if (fcode && L->fcode) memcpy(fcode, L->fcode, sizeof(int) * (size_t)L->m);
    return OR_OK;
}

OR_API int or_amg_dense_inverse(void *p, double *out) {
    amg_t *a = (amg_t *)p;
    if (!a || !a->ainv) return OR_ERR_ARG;
    const int n = a->lv[a->nl - 1].n;
    memcpy(out, a->ainv, sizeof(double) * (size_t)n * n);
    return OR_OK;
}

static void smooth(const level_t *L, const double *r, double *z, double *zn) {
    or_spmv(L->n, L->rowptr, L->col, L->val, z, L->t);
    for (int i = 0; i < L->n; ++i) zn[i] = fma(L->sm[i], r[i] - L->t[i], z[i]);
}

static void vcycle(const amg_t *a, int l, const double *r, double *z) {
    const level_t *L = &a->lv[l];
    if (l == a->nl - 1) {
        if (a->ainv) {
            for (int i = 0; i < L->n; ++i) {
                double s = 0.0;
                for (int j = 0; j < L->n; ++j) s = s + a->ainv[(size_t)i * L->n + j] * r[j];
                z[i] = s;
            }
        } else {
            for (int i = 0; i < L->n; ++i) z[i] = L->sm[i] * r[i];
        }
        return;
    }
    level_t *C = (level_t *)&a->lv[l + 1];
    double *za = L->z, *zb = L->z2;
    for (int i = 0; i < L->n; ++i) za[i] = L->sm[i] * r[i];
    for (int s = 1; s < a->nu; ++s) {
        smooth(L, r, za, zb);
        double *tmp = za; za = zb; zb = tmp;
    }
    or_spmv(L->n, L->rowptr, L->col, L->val, za, L->t);
    for (int i = 0; i < L->n; ++i) L->t[i] = r[i] - L->t[i];
    for (int c = 0; c < C->n; ++c) C->r[c] = 0.0;
    for (int i = 0; i < L->n; ++i) C->r[L->agg[i]] = C->r[L->agg[i]] + L->t[i];
    double *zc = xcalloc((size_t)C->n, sizeof(double));
    vcycle(a, l + 1, C->r, zc);
    for (int i = 0; i < L->n; ++i) za[i] = fma(a->cscale, zc[L->agg[i]], za[i]);
    free(zc);
    for (int s = 0; s < a->nu; ++s) {
        smooth(L, r, za, zb);
        double *tmp = za; za = zb; zb = tmp;
    }
    memcpy(z, za, sizeof(double) * (size_t)L->n);
}

OR_API void or_amg_vcycle(void *p, const double *r, double *z) { vcycle((amg_t *)p, 0, r, z); }

/* FP32 data of the hierarchy (made once) */
static int make32(amg_t *a) {
    if (a->lv[0].val32) return OR_OK;
    for (int l = 0; l < a->nl; ++l) {
        level_t *L = &a->lv[l];
        const int64_t nnz = (int64_t)L->n + 2 * (int64_t)L->m;
        L->val32 = xcalloc((size_t)nnz, sizeof(float));
        L->sm32 = xcalloc((size_t)L->n, sizeof(float));
        L->r32 = xcalloc((size_t)L->n, sizeof(float));
        L->z32 = xcalloc((size_t)L->n, sizeof(float));
        L->z232 = xcalloc((size_t)L->n, sizeof(float));
        L->t32 = xcalloc((size_t)L->n, sizeof(float));
        if (!L->val32 || !L->sm32 || !L->r32 || !L->z32 || !L->z232 || !L->t32)
            return OR_ERR_NOMEM;
        for (int64_t j = 0; j < nnz; ++j) L->val32[j] = (float)L->val[j];
        for (int i = 0; i < L->n; ++i) L->sm32[i] = (float)L->sm[i];
    }
    if (a->ainv) {
        const int n = a->lv[a->nl - 1].n;
        a->ainv32 = xcalloc((size_t)n * n, sizeof(float));
        if (!a->ainv32) return OR_ERR_NOMEM;
        for (int64_t k = 0; k < (int64_t)n * n; ++k) a->ainv32[k] = (float)a->ainv[k];
    }
    return OR_OK;
}

static void smooth32(const level_t *L, const float *r, const float *z, float *zn) {
    or_spmv_f32(L->n, L->rowptr, L->col, L->val32, z, L->t32);
    for (int i = 0; i < L->n; ++i) zn[i] = fmaf(L->sm32[i], r[i] - L->t32[i], z[i]);
}

static void vcycle32(const amg_t *a, int l, const float *r, float *z) {
    const level_t *L = &a->lv[l];
    if (l == a->nl - 1) {
        if (a->ainv32) {
            for (int i = 0; i < L->n; ++i) {
                float s = 0.0f;
                for (int j = 0; j < L->n; ++j) s = s + a->ainv32[(size_t)i * L->n + j] * r[j];
                z[i] = s;
            }
        } else {
            for (int i = 0; i < L->n; ++i) z[i] = L->sm32[i] * r[i];
        }
        return;
    }
    level_t *C = (level_t *)&a->lv[l + 1];
    float *za = L->z32, *zb = L->z232;
    for (int i = 0; i < L->n; ++i) za[i] = L->sm32[i] * r[i];
    for (int s = 1; s < a->nu; ++s) {
        smooth32(L, r, za, zb);
        float *tmp = za; za = zb; zb = tmp;
    }
    or_spmv_f32(L->n, L->rowptr, L->col, L->val32, za, L->t32);
    for (int i = 0; i < L->n; ++i) L->t32[i] = r[i] - L->t32[i];
    for (int c = 0; c < C->n; ++c) C->r32[c] = 0.0f;
    for (int i = 0; i < L->n; ++i) C->r32[L->agg[i]] = C->r32[L->agg[i]] + L->t32[i];
    float *zc = xcalloc((size_t)C->n, sizeof(float));
    vcycle32(a, l + 1, C->r32, zc);
    const float cs = (float)a->cscale;
    for (int i = 0; i < L->n; ++i) za[i] = fmaf(cs, zc[L->agg[i]], za[i]);
    free(zc);
    for (int s = 0; s < a->nu; ++s) {
        smooth32(L, r, za, zb);
        float *tmp = za; za = zb; zb = tmp;
    }
    memcpy(z, za, sizeof(float) * (size_t)L->n);
}

/* the FP32 cycle on a double rhs: (float) r in, (double) z out */
OR_API int or_amg_vcycle32(void *p, const double *r, double *z) {
    amg_t *a = (amg_t *)p;
    if (make32(a) != OR_OK) return OR_ERR_NOMEM;
    const int n = a->lv[0].n;
    float *r32 = xcalloc((size_t)n, sizeof(float)), *z32 = xcalloc((size_t)n, sizeof(float));
    for (int i = 0; i < n; ++i) r32[i] = (float)r[i];
    vcycle32(a, 0, r32, z32);
    for (int i = 0; i < n; ++i) z[i] = (double)z32[i];
    free(r32);
    free(z32);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* PCG                                                                        */
/* ------------------------------------------------------------------------ */
/* DIC: the reciprocal factored diagonal rD (face loop in face order) */
OR_API void or_dic_rd(int n, int m, const int *la, const int *ua, const double *diag,
                      const double *upper, double *rD) {
    for (int i = 0; i < n; ++i) rD[i] = diag[i];
    for (int f = 0; f < m; ++f) rD[ua[f]] -= upper[f] * upper[f] / rD[la[f]];
    for (int i = 0; i < n; ++i) rD[i] = 1.0 / rD[i];
}

/* DIC: z = M^-1 r (forward sweep in face order, backward in reverse face order) */
OR_API void or_dic_apply(int n, int m, const int *la, const int *ua, const double *rD,
                         const double *upper, const double *r, double *z) {
    for (int i = 0; i < n; ++i) z[i] = rD[i] * r[i];
    for (int f = 0; f < m; ++f) z[ua[f]] -= rD[ua[f]] * upper[f] * z[la[f]];
    for (int f = m - 1; f >= 0; --f) z[la[f]] -= rD[la[f]] * upper[f] * z[ua[f]];
}

/* DILU (OpenFOAM's DILUPreconditioner, published algorithm; for asymmetric
 * matrices): M = (D* + L) D*^-1 (D* + U) with D* chosen so that diag(M) = diag(A):
 *   rD = diag; for f in face order: rD[u] -= upper[f] lower[f] / rD[l]; rD = 1 / rD;
 *   z = rD r; for f in face order: z[u] -= rD[u] lower[f] z[l];
 *   for f in reverse face order: z[l] -= rD[l] upper[f] z[u].
 * transpose: z = M^-T r, the same sweeps with lower and upper exchanged.  With
 * lower == upper every operation is DIC's, so the bits are DIC's. */
OR_API void or_dilu_rd(int n, int m, const int *la, const int *ua, const double *diag,
                       const double *lower, const double *upper, double *rD) {
    for (int i = 0; i < n; ++i) rD[i] = diag[i];
    for (int f = 0; f < m; ++f) rD[ua[f]] -= upper[f] * lower[f] / rD[la[f]];
    for (int i = 0; i < n; ++i) rD[i] = 1.0 / rD[i];
}

OR_API void or_dilu_apply(int n, int m, const int *la, const int *ua, const double *rD,
                          const double *lower, const double *upper, int transpose,
                          const double *r, double *z) {
    const double *cf = transpose ? upper : lower, *cb = transpose ? lower : upper;
    for (int i = 0; i < n; ++i) z[i] = rD[i] * r[i];
    for (int f = 0; f < m; ++f) z[ua[f]] -= rD[ua[f]] * cf[f] * z[la[f]];
    for (int f = m - 1; f >= 0; --f) z[la[f]] -= rD[la[f]] * cb[f] * z[ua[f]];
}

typedef struct {
    int m;
    const int *la, *ua;
    const double *upper;
    double *rD;
} dic_t;

static void precond(int pc, const amg_t *amg, const dic_t *dic, int n, const double *dinv,
                    const double *r, double *z) {
    if (pc == PC_DIC) {
        or_dic_apply(n, dic->m, dic->la, dic->ua, dic->rD, dic->upper, r, z);
    } else if (pc == PC_NONE) {
        memcpy(z, r, sizeof(double) * (size_t)n);
    } else if (pc == PC_DIAGONAL) {
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    } else if (pc == PC_GAMG) {
        vcycle(amg, 0, r, z);
    } else {
        (void)or_amg_vcycle32((void *)amg, r, z);
    }
}

/*
 * PCG on an LDU system (x holds x0).  pc: 0 none, 1 diagonal, 2 GAMG with
 * (omega, nu, cscale).  error / max_steps (<= 0 unbounded) / rel_tol /
 * min_steps as or_solve_ctl.  hist as or_solve; returns 0, or an error code.
 */
OR_API int or_pcg(int pc, int N, int M, const int *la, const int *ua, const double *diag,
                  const double *lower, const double *upper, const double *b, double *x,
                  double error, int max_steps, double rel_tol, int64_t min_steps, double omega,
                  int nu, double cscale, double *hist, int hist_cap, int *hist_len) {
    if (pc < PC_NONE || pc > PC_DIC || N < 1) return OR_ERR_ARG;
    amg_t *amg = NULL;
    level_t A;
    memset(&A, 0, sizeof(A));
    A.n = N;
    A.m = M;
    A.la = (int *)la;
    A.ua = (int *)ua;
    A.diag = (double *)diag;
    A.lower = (double *)lower;
    A.upper = (double *)upper;
    if (level_csr(&A) != OR_OK) return OR_ERR_ARG;
    or_ldu2csr_values(N, M, (int64_t)N + 2 * (int64_t)M, A.perm, diag, lower, upper, A.val);
    if (pc >= PC_GAMG) {
        amg = or_amg_build(N, M, la, ua, diag, lower, upper, omega, nu, cscale);
        if (!amg || (pc == PC_GAMG32 && make32(amg) != OR_OK)) {
            or_amg_free(amg);
            free(A.rowptr); free(A.col); free(A.perm); free(A.val);
            return OR_ERR_NOMEM;
        }
    }
    double *v = xcalloc((size_t)N * 5, sizeof(double));
    double *r = v, *z = v + N, *p = v + 2 * (size_t)N, *q = v + 3 * (size_t)N;
    double *dinv = v + 4 * (size_t)N;
    for (int i = 0; i < N; ++i) dinv[i] = 1.0 / diag[i];
    dic_t dic = {M, la, ua, upper, NULL};
    if (pc == PC_DIC) {
        dic.rD = xcalloc((size_t)N, sizeof(double));
        or_dic_rd(N, M, la, ua, diag, upper, dic.rD);
    }
    or_spmv(N, A.rowptr, A.col, A.val, x, q);
    for (int i = 0; i < N; ++i) r[i] = b[i] - q[i];
    double norm = sqrt(or_dot(N, r, r));
    int len = 0;
    if (hist && len < hist_cap) hist[len] = norm;
    len++;
    if (rel_tol > 0.0) error = fmax(error, rel_tol * norm);
    const int64_t ms = max_steps <= 0 ? INT32_MAX : max_steps;
    double rho = 0.0, beta = 0.0;
    if ((norm > error) || min_steps > 0) {
        precond(pc, amg, &dic, N, dinv, r, z);
        rho = or_dot(N, r, z);
    }
    for (int64_t k = 1; (k <= ms && norm > error) || k <= min_steps; ++k) {
        for (int i = 0; i < N; ++i) p[i] = (k == 1) ? z[i] : fma(beta, p[i], z[i]);
        or_spmv(N, A.rowptr, A.col, A.val, p, q);
        const double alpha = rho / or_dot(N, p, q);
        for (int i = 0; i < N; ++i) {
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, q[i], r[i]);
        }
        norm = sqrt(or_dot(N, r, r));
        if (hist && len < hist_cap) hist[len] = norm;
        len++;
        if (!((k + 1 <= ms && norm > error) || k + 1 <= min_steps)) break;
        const double gamma = pc == PC_GAMG32 ? or_dot(N, r, z) : 0.0;
        precond(pc, amg, &dic, N, dinv, r, z);
        const double rho_new = or_dot(N, r, z);
        beta = pc == PC_GAMG32 ? (rho_new - gamma) / rho : rho_new / rho;
        rho = rho_new;
    }
    if (hist_len) *hist_len = len;
    free(v);
    free(dic.rD);
    free(A.rowptr); free(A.col); free(A.perm); free(A.val);
    or_amg_free(amg);
    return OR_OK;
}

/*
 * BiCG preconditioned by M on an LDU system (no counterpart in the reference;
 * the loop, breakdown test, stopping rule and history of gis_oracle.c's bicg,
 * SolverSingleGPU.h:568-675): z = M^-1 r, z* = M^-T r*, rho = (z, r*),
 * p = z or fma(theta, p, z), p* likewise from z*.  pc: PC_DIAGONAL (M = D) or
 * PC_DILU.  A^T is A's pattern with lower and upper exchanged.  error /
 * max_steps (<= 0 unbounded) / int_bound / rel_tol / min_steps as
 * or_solve_ctl; returns 0, 10 (breakdown) or an error code.
 */
OR_API int or_pbicg(int pc, int N, int M, const int *la, const int *ua, const double *diag,
                    const double *lower, const double *upper, const double *b, double *x,
                    double error, int max_steps, double int_bound, double rel_tol,
                    int64_t min_steps, double *hist, int hist_cap, int *hist_len) {
    if ((pc != PC_DIAGONAL && pc != PC_DILU) || N < 1) return OR_ERR_ARG;
    level_t A;
    memset(&A, 0, sizeof(A));
    A.n = N;
    A.m = M;
    A.la = (int *)la;
    A.ua = (int *)ua;
    if (level_csr(&A) != OR_OK) return OR_ERR_ARG;
    const int64_t nnz = (int64_t)N + 2 * (int64_t)M;
    double *valT = xcalloc((size_t)nnz, sizeof(double));
    double *v = xcalloc((size_t)N * 9, sizeof(double));
    if (!valT || !v) {
        free(valT); free(v); free(A.rowptr); free(A.col); free(A.perm); free(A.val);
        return OR_ERR_NOMEM;
    }
    or_ldu2csr_values(N, M, nnz, A.perm, diag, lower, upper, A.val);
    or_ldu2csr_values(N, M, nnz, A.perm, diag, upper, lower, valT);
    double *r = v, *rs = v + N, *p = v + 2 * (size_t)N, *ps = v + 3 * (size_t)N;
    double *q = v + 4 * (size_t)N, *qs = v + 5 * (size_t)N, *z = v + 6 * (size_t)N;
    double *zs = v + 7 * (size_t)N, *rD = v + 8 * (size_t)N;
    if (pc == PC_DILU) {
        or_dilu_rd(N, M, la, ua, diag, lower, upper, rD);
    } else {
        for (int i = 0; i < N; ++i) rD[i] = 1.0 / diag[i];
    }
#define PRECOND(src, dst, T)                                                      \
    do {                                                                          \
        if (pc == PC_DILU)                                                        \
            or_dilu_apply(N, M, la, ua, rD, lower, upper, (T), (src), (dst));     \
        else                                                                      \
            for (int i_ = 0; i_ < N; ++i_) (dst)[i_] = rD[i_] * (src)[i_];        \
    } while (0)
    or_spmv(N, A.rowptr, A.col, A.val, x, q);
    for (int i = 0; i < N; ++i) r[i] = b[i] - q[i];
    memcpy(rs, r, sizeof(double) * (size_t)N);
    PRECOND(r, z, 0);
    PRECOND(rs, zs, 1);
    memcpy(p, z, sizeof(double) * (size_t)N);
    memcpy(ps, zs, sizeof(double) * (size_t)N);
    double rho0 = or_dot(N, z, rs), rho1 = 0.0;
    double norm = sqrt(or_dot(N, r, r));
    int len = 0;
    if (hist && len < hist_cap) hist[len] = norm;
    len++;
    if (rel_tol > 0.0) error = fmax(error, rel_tol * norm);
    const int64_t ms = max_steps <= 0 ? INT32_MAX : max_steps;
    int status = OR_OK;
    for (int64_t k = 1; k <= ms || k <= min_steps; ++k) {
        or_spmv(N, A.rowptr, A.col, A.val, p, q);
        const double xi = rho0 / or_dot(N, ps, q);
        if (xi > -int_bound && xi < int_bound) {
            status = OR_BREAKDOWN_STATUS;
            break;
        }
        for (int i = 0; i < N; ++i) {
            x[i] = fma(xi, p[i], x[i]);
            r[i] = fma(-xi, q[i], r[i]);
        }
        norm = sqrt(or_dot(N, r, r));
        if (hist && len < hist_cap) hist[len] = norm;
        len++;
        if (norm < error && k >= min_steps) break;
        or_spmv(N, A.rowptr, A.col, valT, ps, qs);
        for (int i = 0; i < N; ++i) rs[i] = fma(-xi, qs[i], rs[i]);
        PRECOND(r, z, 0);
        rho1 = rho0;
        rho0 = or_dot(N, z, rs);
        const double theta = rho0 / rho1;
        for (int i = 0; i < N; ++i) p[i] = fma(theta, p[i], z[i]);
        PRECOND(rs, zs, 1);
        for (int i = 0; i < N; ++i) ps[i] = fma(theta, ps[i], zs[i]);
    }
#undef PRECOND
    if (hist_len) *hist_len = len;
    free(v);
    free(valT);
    free(A.rowptr); free(A.col); free(A.perm); free(A.val);
    return status;
}
