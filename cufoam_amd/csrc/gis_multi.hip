// Several right-hand sides that share one matrix (the three velocity
// components of one OpenFOAM momentum matrix, or a batch of adjoint loads):
// one SpMM launch streams the matrix once for every column.  Each column keeps
// its own scalars and stopping rule in its own KState and its reductions use
// the single-RHS order, so column k is bit-identical to a single-RHS solve of
// B[k] (CG :366-445, BiCG :568-675 of gis/SparseSystem/SolverSingleGPU.h).
//
// The adjoint of x = A^{-1} b for losses L_k = w_k . x:  A^T lambda_k = w_k,
// dL_k/db = lambda_k, dL_k/dA_ij = -lambda_k[i] x[j] on the pattern, written
// straight into the LDU coefficient slots through the CSR value permutation.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "gis_kernels.hpp"

namespace gis {

namespace {

constexpr int kMaxRhs = 8;

// Every kernel below is templated on the Krylov value type T: double for the
// FP64 solvers, float for the inner loop of the mixed-precision refinement.
// Products, row sums and vector updates run in T (fused multiply-adds only as
// fmat); every reduction accumulates in double, so T = double is exactly the
// FP64 arithmetic and T = float that of the oracle's or_mixed_krylov.
__device__ __forceinline__ double fmat(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmat(float a, float b, float c) { return fmaf(a, b, c); }

// y[c][k] for the four 64-row slices of this wave, columns with bit k set in
// `active`.  A slice with a column-offset table (m.ooff set: the handle's
// schedule is SellD) reads 64 values per table entry and the row masks; the
// others read the int32 slot columns.  Either way one value load feeds the
// gathers of every column, the loop has no per-slot branch (a clear mask bit
// or a padding slot gathers a valid row and leaves the sums untouched through
// a select), and each row sums its entries in storage order with rounded
// products: the bits of tile_sell / tile_sell_d, column by column.
template <int NR, typename T, class XOp>
__device__ __forceinline__ void sell_multi(const Mat<T>& m, int N, int row0, XOp&& xop,
                                           unsigned active, T (&y)[4][NR]) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
    const int last = N - 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int k = 0; k < NR; ++k) y[c][k] = T(0);
        const int s = s0 + c;
        if (s >= nslices) continue;
        const int row = s * kSlice + lane;
        const int o = m.ooff ? m.ooff[s] : 0;
        const int U = m.ooff ? m.ooff[s + 1] - o : 0;
        if (U > 0) {
            const unsigned mask = row_mask(m.dmask, m.dmask8, U, row, N);
            const T* dv = m.dval + (size_t)o * kSlice + lane;
            const int* offs = m.doffs + o;
            // a structured slice (at most 8 table entries) in one batch: every
            // value load and gather of the slice in flight together
            for (int j = 0; j < U; j += 8) {
                T v[8];
                int o8[8], cc[8];
                // the offsets as one block of scalar loads before any use: loaded
                // one by one between the gathers, each waited for its own load
#pragma unroll
                for (int u = 0; u < 8; ++u) o8[u] = offs[min(j + u, U - 1)];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    v[u] = __builtin_nontemporal_load(dv + (size_t)min(j + u, U - 1) * kSlice);
#pragma unroll
                for (int u = 0; u < 8; ++u) cc[u] = min(max(row + o8[u], 0), last);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool on = (mask >> (j + u)) & 1u;
#pragma unroll
                    for (int k = 0; k < NR; ++k) {
                        const bool use = on && ((active >> k) & 1u);
                        const T t = y[c][k] + v[u] * xop(k, cc[u]);
                        y[c][k] = use ? t : y[c][k];
                    }
                }
            }
        } else {
            const int a = m.soff[s];
            const int len = (m.soff[s + 1] - a) / kSlice;
            const int safe = min(row, last);
            for (int j = 0; j < len; j += 4) {
                int cc[4];
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int idx = a + min(j + u, len - 1) * kSlice + lane;
                    cc[u] = __builtin_nontemporal_load(m.scol + idx);
                    v[u] = __builtin_nontemporal_load(m.sval + idx);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool on = (j + u < len) && cc[u] >= 0;
                    const int col = on ? cc[u] : safe;
#pragma unroll
                    for (int k = 0; k < NR; ++k) {
                        const bool use = on && ((active >> k) & 1u);
                        const T t = y[c][k] + v[u] * xop(k, col);
                        y[c][k] = use ? t : y[c][k];
                    }
                }
            }
        }
    }
}

// The matrix view the multi-RHS kernels read: the offset tables only when the
// handle's SpMV schedule for that precision is SellD (sell_multi then reads
// the tabled slices through them), else the int32 slot columns throughout.
template <typename T>
Mat<T> multi_view(Mat<T> m, SpmvKind kind) {
    if (kind != SpmvKind::SellD) m.ooff = nullptr;
    return m;
}

// column k of the wave's slices -> this thread's kLeaf consecutive rows
template <int NR, typename T>
__device__ __forceinline__ void to_leaf(const T (&y)[4][NR], int k, T* wave_lds, T (&o)[4]) {
    const int lane = threadIdx.x & 63;
    wave_sync_lds();
#pragma unroll
    for (int c = 0; c < 4; ++c) wave_lds[c * kSlice + lane] = y[c][k];
    wave_sync_lds();
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = wave_lds[lane * kLeaf + q];
}

template <int NR>
__device__ __forceinline__ unsigned active_mask(const KState* st) {
    unsigned a = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k)
        if (!st[k].done) a |= 1u << k;
    return a;
}

// r_k = b_k - A x_k, optional copy, rho0 = (r, r), norm0, hist[0].  Columns in
// `skip` are left alone and marked done.
template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_msetup(Mat<T> m, int N, const T* __restrict__ B,
                                                const T* __restrict__ X, T* __restrict__ R,
                                                T* __restrict__ R2, KState* st, RedWs red,
                                                int done_rule, unsigned skip) {
    __shared__ T sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    const unsigned act = ((1u << NR) - 1) & ~skip;
    T y[4][NR];
    sell_multi<NR, T>(m, N, t.row0, [&](int k, int c) { return X[k * n + c]; }, act, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        T yl[4], bv[4], rv[4];
        to_leaf<NR, T>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        load4(B + k * n, t.i0, N, bv);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rv[q] = bv[q] - yl[q];
            if (t.i0 + q < N) leaf[k] = fma((double)rv[q], (double)rv[q], leaf[k]);
        }
        store4(R + k * n, t.i0, N, rv);
        if (R2) store4(R2 + k * n, t.i0, N, rv);
    }
    finish_tile_slots<NR>(red, st, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(act & (1u << k))) {
                st[k].done = 1;
                continue;
            }
            const double norm = sqrt(v[k]);
            st[k].rho0 = v[k];
            st[k].rho1 = 0.0;
            st[k].norm = norm;
            st[k].norm0 = norm;
            st[k].k = 0;
            st[k].status = kOK;
            if (st[k].hist_cap > 0) st[k].hist[0] = norm;
            st[k].done = (done_rule == 0) ? !(norm > st[k].error) : 0;
        }
    }, NR);
}

// ---- directions (both methods), streamed ----------------------------------------
// p_k of iteration kk lives in buffer kk & 1.  CG: p = r + beta p_old and the
// deferred x += alpha_prev p_old (SolverSingleGPU.h:366-445).  BiCG: p = r +
// beta p_old, p* = r* + beta p* (SolverSingleGPU.h:568-675).  The SpMM that
// follows then gathers one vector per column instead of forming p at every
// gathered column from two.
template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mdir(int N, const T* __restrict__ R,
                                              const T* __restrict__ RS, T* __restrict__ P0,
                                              T* __restrict__ P1, T* __restrict__ PS,
                                              T* __restrict__ X, const KState* st, int bicg) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    StreamRows t(N);
    const size_t n = (size_t)N;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (!(act & (1u << k))) continue;
        const int kk = st[k].k + 1;
        const T beta = (T)st[k].beta;
        const T alpha_prev = (T)st[k].alpha;
        const T* pold = ((kk & 1) ? P0 : P1) + k * n;
        T* pnew = ((kk & 1) ? P1 : P0) + k * n;
        T rv[4];
        load4(R + k * n, t.i0, N, rv);
        if (kk == 1) {
            store4(pnew, t.i0, N, rv);
            if (bicg) {
                T rsv[4];
                load4(RS + k * n, t.i0, N, rsv);
                store4(PS + k * n, t.i0, N, rsv);
            }
            continue;
        }
        T pv[4], pn[4];
        load4(pold, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q) pn[q] = fmat(beta, pv[q], rv[q]);
        store4(pnew, t.i0, N, pn);
        if (bicg) {
            T rsv[4], psv[4];
            load4(RS + k * n, t.i0, N, rsv);
            load4(PS + k * n, t.i0, N, psv);
#pragma unroll
            for (int q = 0; q < 4; ++q) psv[q] = fmat(beta, psv[q], rsv[q]);
            store4(PS + k * n, t.i0, N, psv);
        } else {
            T xv[4];
            load4(X + k * n, t.i0, N, xv);
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = fmat(alpha_prev, pv[q], xv[q]);
            store4(X + k * n, t.i0, N, xv);
        }
    }
}

// ---- CG ---------------------------------------------------------------------
// after k_mdir: q = A p_k, (p_k, q) -> alpha_k
template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mcg1(Mat<T> m, int N, const T* __restrict__ P0,
                                              const T* __restrict__ P1, T* __restrict__ Q,
                                              KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ T sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    // columns advance in lockstep until they stop: every active column is at
    // the same iteration, hence in the same buffer
    const T* P = ((st[__ffs(act) - 1].k + 1) & 1) ? P1 : P0;
    T y[4][NR];
    sell_multi<NR, T>(m, N, t.row0, [&](int k, int c) { return P[k * n + c]; }, act, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        T yl[4];
        to_leaf<NR, T>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        T pv[4];
        load4(P + k * n, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t.i0 + q < N) leaf[k] = fma((double)pv[q], (double)yl[q], leaf[k]);
        store4(Q + k * n, t.i0, N, yl);
    }
    finish_tile_slots<NR>(red, st, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (act & (1u << k)) st[k].alpha = st[k].rho0 / v[k];
    }, NR);
}

template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mcg2(int N, T* __restrict__ R, const T* __restrict__ Q,
                                              KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ RedShared shr;
    StreamRows t(N);
    const size_t n = (size_t)N;
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        const T alpha = (T)st[k].alpha;
        T rv[4], qv[4];
        load4(R + k * n, t.i0, N, rv);
        load4(Q + k * n, t.i0, N, qv);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rv[q] = fmat(-alpha, qv[q], rv[q]);
            if (t.i0 + q < N) leaf[k] = fma((double)rv[q], (double)rv[q], leaf[k]);
        }
        store4(R + k * n, t.i0, N, rv);
    }
    finish_tile_slots<NR>(red, st, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(act & (1u << k))) continue;
            KState* s = &st[k];
            s->rho1 = s->rho0;
            s->rho0 = v[k];
            s->beta = s->rho0 / s->rho1;
            const double norm = sqrt(v[k]);
            record_norm(s, norm);
            if (!(norm > s->error) || s->k >= s->max_steps) s->done = 1;
        }
    }, NR);
}

template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mcg_flush(int N, T* __restrict__ X,
                                                   const T* __restrict__ P0,
                                                   const T* __restrict__ P1, const KState* st,
                                                   unsigned skip) {
    StreamRows t(N);
    const size_t n = (size_t)N;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int it = st[k].k;
        if (it < 1 || (skip & (1u << k))) continue;
        const T alpha = (T)st[k].alpha;
        const T* p = ((it & 1) ? P1 : P0) + k * n;
        T xv[4], pv[4];
        load4(X + k * n, t.i0, N, xv);
        load4(p, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = fmat(alpha, pv[q], xv[q]);
        store4(X + k * n, t.i0, N, xv);
    }
}

// ---- BiCG (the matrix m is "A", mt is its transpose) -------------------------
// after k_mdir: q = A p_k, (p*_k, q) -> xi = alpha_k, breakdown test
template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mbi1(Mat<T> m, int N, const T* __restrict__ P0,
                                              const T* __restrict__ P1, const T* __restrict__ PS,
                                              T* __restrict__ Q, KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ T sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    const T* P = ((st[__ffs(act) - 1].k + 1) & 1) ? P1 : P0;
    T y[4][NR];
    sell_multi<NR, T>(m, N, t.row0, [&](int k, int c) { return P[k * n + c]; }, act, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        T yl[4];
        to_leaf<NR, T>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        T psv[4];
        load4(PS + k * n, t.i0, N, psv);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t.i0 + q < N) leaf[k] = fma((double)psv[q], (double)yl[q], leaf[k]);
        store4(Q + k * n, t.i0, N, yl);
    }
    finish_tile_slots<NR>(red, st, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(act & (1u << k))) continue;
            KState* s = &st[k];
            const double xi = s->rho0 / v[k];
            s->alpha = xi;
            if (xi > -s->int_bound && xi < s->int_bound) {
                s->status = kBreakdown;
                s->done = 1;
            }
        }
    }, NR);
}

// q* = A^T p*_k; x += xi p, r -= xi q, r* -= xi q*, ||r||, (r, r*) -> beta, stop
template <int NR, typename T>
__global__ __launch_bounds__(kWG) void k_mbi2(Mat<T> mt, int N, T* __restrict__ R,
                                              T* __restrict__ RS, const T* __restrict__ P0,
                                              const T* __restrict__ P1, const T* __restrict__ PS,
                                              const T* __restrict__ Q, T* __restrict__ X,
                                              KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ T sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    T y[4][NR];
    sell_multi<NR, T>(mt, N, t.row0, [&](int k, int c) { return PS[k * n + c]; }, act, y);
    double leaf[2 * NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        T yl[4];
        to_leaf<NR, T>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[2 * k] = 0.0;
        leaf[2 * k + 1] = 0.0;
        if (!(act & (1u << k))) continue;
        const T xi = (T)st[k].alpha;
        const T* p = (((st[k].k + 1) & 1) ? P1 : P0) + k * n;
        T xv[4], pv[4], rv[4], qv[4], rsv[4];
        load4(X + k * n, t.i0, N, xv);
        load4(p, t.i0, N, pv);
        load4(R + k * n, t.i0, N, rv);
        load4(Q + k * n, t.i0, N, qv);
        load4(RS + k * n, t.i0, N, rsv);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xv[q] = fmat(xi, pv[q], xv[q]);
            rv[q] = fmat(-xi, qv[q], rv[q]);
            rsv[q] = fmat(-xi, yl[q], rsv[q]);
            if (t.i0 + q < N) {
                leaf[2 * k] = fma((double)rv[q], (double)rv[q], leaf[2 * k]);
                leaf[2 * k + 1] = fma((double)rv[q], (double)rsv[q], leaf[2 * k + 1]);
            }
        }
        store4(X + k * n, t.i0, N, xv);
        store4(R + k * n, t.i0, N, rv);
        store4(RS + k * n, t.i0, N, rsv);
    }
    finish_tile_slots<2 * NR>(red, st, t.tile, leaf, shr, [&](const double (&v)[2 * NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(act & (1u << k))) continue;
            KState* s = &st[k];
            const double norm = sqrt(v[2 * k]);
            record_norm(s, norm);
            if (norm < s->error || s->k >= s->max_steps) {
                s->done = 1;
            } else {
                s->rho1 = s->rho0;
                s->rho0 = v[2 * k + 1];
                s->beta = s->rho0 / s->rho1;
            }
        }
    }, NR);
}

// ---- mixed-precision refinement: the FP64 outer steps ------------------------
// R64 = B - A X in FP64 (never stored), R32 = (float) R64, ||R64||, and on the
// first step ||B||; per column outer state in `ost` (norm, error = rel_tol ||b||,
// done, k = outer steps, hist = outer residual norms).
template <int NR>
__global__ __launch_bounds__(kWG) void k_mres(Mat<double> m, int N, const double* __restrict__ B,
                                              const double* __restrict__ X,
                                              float* __restrict__ R32, KState* ost, RedWs red,
                                              double rel_tol, int first) {
    __shared__ double sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    const unsigned act = first ? (1u << NR) - 1 : active_mask<NR>(ost);
    if (!act) return;
    double y[4][NR];
    sell_multi<NR, double>(m, N, t.row0, [&](int k, int c) { return X[k * n + c]; }, act, y);
    double leaf[2 * NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double yl[4], bv[4];
        to_leaf<NR, double>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[2 * k] = 0.0;
        leaf[2 * k + 1] = 0.0;
        if (!(act & (1u << k))) continue;
        load4(B + k * n, t.i0, N, bv);
        float rf[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double r = bv[q] - yl[q];
            rf[q] = (float)r;
            if (t.i0 + q < N) {
                leaf[2 * k] = fma(r, r, leaf[2 * k]);
                leaf[2 * k + 1] = fma(bv[q], bv[q], leaf[2 * k + 1]);
            }
        }
        store4(R32 + k * n, t.i0, N, rf);
    }
    finish_tile_slots<2 * NR>(red, ost, t.tile, leaf, shr, [&](const double (&v)[2 * NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(act & (1u << k))) continue;
            KState* s = &ost[k];
            if (first) {
                s->aux[3] = sqrt(v[2 * k + 1]);
                s->error = rel_tol * s->aux[3];
                s->k = 0;
                s->nhist = 0;
                s->status = kOK;
                s->done = 0;
                s->aux[0] = 0.0;     // inner iterations, summed
            }
            const double rn = sqrt(v[2 * k]);
            s->norm = rn;
            if (first) s->norm0 = rn;
            if (s->nhist < s->hist_cap) s->hist[s->nhist] = rn;
            s->nhist = s->nhist + 1;
            if (rn != rn) {
                s->status = kNaN;
                s->done = 1;
            } else if (!(rn > s->error)) {
                s->done = 1;
            } else if (s->k >= s->max_steps) {
                s->status = kNotConverged;
                s->done = 1;
            }
        }
    }, NR);
}

// X += (double) D for the columns refined in this outer step; counts the step
template <int NR>
__global__ __launch_bounds__(kWG) void k_mcorr(int N, double* __restrict__ X,
                                               const float* __restrict__ D,
                                               const KState* ist, KState* ost, unsigned refined) {
    StreamRows t(N);
    const size_t n = (size_t)N;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (!(refined & (1u << k))) continue;
        double xv[4];
        float dv[4];
        load4(X + k * n, t.i0, N, xv);
        load4(D + k * n, t.i0, N, dv);
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = xv[q] + (double)dv[q];
        store4(X + k * n, t.i0, N, xv);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!(refined & (1u << k))) continue;
            ost[k].k = ost[k].k + 1;
            ost[k].aux[0] = ost[k].aux[0] + (double)ist[k].k;
        }
    }
}

// dL_k/dA_ij = -lambda_k[i] x[j] on the CSR pattern, folded into the LDU slots
// through the value permutation (each slot is written exactly once).
__global__ __launch_bounds__(kWG) void k_sddmm_fold(int N, int M, int nrhs,
                                                    const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const int* __restrict__ perm,
                                                    const double* __restrict__ Lam,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ d_diag,
                                                    double* __restrict__ d_lower,
                                                    double* __restrict__ d_upper) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= N) return;
    const size_t n = (size_t)N, mm = (size_t)M;
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const int j = col[p];
        const int code = perm[p];
        const double xj = x[j];
        for (int k = 0; k < nrhs; ++k) {
            const double g = -(Lam[k * n + i] * xj);
            if (code < M) d_lower[k * mm + code] = g;
            else if (code < M + N) d_diag[k * n + (code - M)] = g;
            else if (code < 2 * M + N) d_upper[k * mm + (code - M - N)] = g;
            // coupling entries (codes >= 2M + N) have no LDU slot to fold into
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct MultiBufs {
    KState* st = nullptr;       // device, nrhs states
    KState* host = nullptr;     // pinned mirror
};

// states: h->mstate holds 2 kMaxRhs KStates (the Krylov states, then the outer
// states of the mixed-precision refinement), mirrored in pinned host memory
int ensure_multi_states(gis_context* h) {
    GIS_HIP(h, h->mstate.ensure(sizeof(KState) * 2 * kMaxRhs));
    if (!h->mstate_host) {
        void* p = nullptr;
        GIS_HIP(h, hipHostMalloc(&p, sizeof(KState) * 2 * kMaxRhs, hipHostMallocDefault));
        h->mstate_host = static_cast<KState*>(p);
    }
    return kOK;
}

int multi_prepare(gis_context* h, int nrhs, const SolveArgs& a, size_t vec_bytes,
                  MultiBufs& mb) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    GIS_HIP(h, h->vec.ensure(vec_bytes));
    const int cap = std::max(1, a.hist_cap);
    GIS_HIP(h, h->hist.ensure(sizeof(double) * (size_t)cap * nrhs));
    GIS_TRY(ensure_multi_states(h));
    mb.st = h->mstate.as<KState>();
    mb.host = h->mstate_host;
    for (int k = 0; k < nrhs; ++k) {
        KState s{};
        s.error = a.error;
        s.max_steps = a.max_steps <= 0 ? (long long)INT_MAX : a.max_steps;
        s.slot_fault = h->slot_fault;
        s.int_bound = a.int_bound;
        s.hist = a.hist_cap > 0 ? h->hist.as<double>() + (size_t)k * cap : nullptr;
        s.hist_cap = a.hist_cap > 0 ? cap : 0;
        mb.host[k] = s;
    }
    GIS_HIP(h, hipMemcpyAsync(mb.st, mb.host, sizeof(KState) * nrhs, hipMemcpyHostToDevice,
                              h->stream));
    return kOK;
}

int multi_read(gis_context* h, int nrhs, MultiBufs& mb, bool* all_done) {
    GIS_HIP(h, hipMemcpyAsync(mb.host, mb.st, sizeof(KState) * nrhs, hipMemcpyDeviceToHost,
                              h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    GIS_TRY(timer_collect(h));
    bool d = true;
    for (int k = 0; k < nrhs; ++k) {
        d = d && mb.host[k].done;
        // a reduction whose wait ran out: rebuild the workspace (and its slots)
        if (mb.host[k].status == kErrSync) h->red_rows = -1;
    }
    *all_done = d;
    return kOK;
}

// The Krylov loop of a multi-RHS CG (method 0) or BiCG (method 1) in value type
// T on the workspace `base` (6 NR N values of T: R, P0, P1, Q, R*, P*); columns
// in `skip` take no part.
template <int NR, typename T>
int run_multi(gis_context* h, int method, const Mat<T>& mA, const Mat<T>& mAt,
              long long max_steps, T* X, const T* B, T* base, MultiBufs& mb, unsigned skip) {
    const int N = h->N;
    const size_t n = (size_t)N;
    const dim3 g = tiles_grid(N);
    T* R = base;
    T* P0 = base + 1 * n * NR;
    T* P1 = base + 2 * n * NR;
    T* Q = base + 3 * n * NR;
    T* RS = base + 4 * n * NR;
    T* PS = base + 5 * n * NR;
    const bool bicg = method == 1;
    hipLaunchKernelGGL((k_msetup<NR, T>), g, dim3(kWG), 0, h->stream, mA, N, B, X, R,
                       bicg ? RS : nullptr, mb.st, h->red, bicg ? 1 : 0, skip);
    GIS_HIP(h, hipGetLastError());
    const long long ms = max_steps <= 0 ? (long long)INT_MAX : max_steps;
    long long launched = 0;
    int block = 4;
    for (;;) {
        for (int i = 0; i < block && launched < ms; ++i, ++launched) {
            GIS_TRY(timer_begin(h, 4));
            hipLaunchKernelGGL((k_mdir<NR, T>), g, dim3(kWG), 0, h->stream, N, R, RS, P0, P1,
                               PS, X, mb.st, bicg ? 1 : 0);
            GIS_TRY(timer_end(h));
            GIS_TRY(timer_begin(h, 0));
            if (bicg)
                hipLaunchKernelGGL((k_mbi1<NR, T>), g, dim3(kWG), 0, h->stream, mA, N, P0, P1,
                                   PS, Q, mb.st, h->red);
            else
                hipLaunchKernelGGL((k_mcg1<NR, T>), g, dim3(kWG), 0, h->stream, mA, N, P0, P1,
                                   Q, mb.st, h->red);
            GIS_TRY(timer_end(h));
            GIS_TRY(timer_begin(h, 1));
            if (bicg)
                hipLaunchKernelGGL((k_mbi2<NR, T>), g, dim3(kWG), 0, h->stream, mAt, N, R, RS,
                                   P0, P1, PS, Q, X, mb.st, h->red);
            else
                hipLaunchKernelGGL((k_mcg2<NR, T>), g, dim3(kWG), 0, h->stream, N, R, Q, mb.st,
                                   h->red);
            GIS_TRY(timer_end(h));
        }
        GIS_HIP(h, hipGetLastError());
        bool done = false;
        GIS_TRY(multi_read(h, NR, mb, &done));
        if (done || launched >= ms) break;
        block = std::min(block * 2, std::max(1, h->sched.block_iters));
    }
    if (!bicg) {
        hipLaunchKernelGGL((k_mcg_flush<NR, T>), g, dim3(kWG), 0, h->stream, N, X, P0, P1,
                           mb.st, skip);
        GIS_HIP(h, hipGetLastError());
    }
    return kOK;
}

template <int NR>
int solve_multi_nr(gis_context* h, int method, bool transpose, const SolveArgs& a, double* X,
                   const double* B, MultiBufs& mb) {
    const Mat<double> mA = multi_view(transpose ? mat64T(h) : mat64(h), h->sched.spmv64);
    const Mat<double> mAt = multi_view(transpose ? mat64(h) : mat64T(h), h->sched.spmv64);
    return run_multi<NR, double>(h, method, mA, mAt, a.max_steps, X, B, h->vec.as<double>(), mb,
                                 0u);
}

// Restarted refinement: FP64 residual, FP32 Krylov solve for the correction,
// FP64 correction, per column until ||b - A x|| <= rel_tol ||b|| or max_outer
// corrections.
template <int NR>
int solve_multi_mixed_nr(gis_context* h, int method, const MultiMixedArgs& a, double* X,
                         const double* B, MultiBufs& mb, KState* ost, KState* ost_host) {
    const int N = h->N;
    const size_t n = (size_t)N;
    const dim3 g = tiles_grid(N);
    const Mat<double> m64 = multi_view(mat64(h), h->sched.spmv64);
    const Mat<float> m32 = multi_view(mat32(h), h->sched.spmv32);
    const Mat<float> m32T = multi_view(mat32T(h), h->sched.spmv32);
    float* base = h->vec.as<float>();
    float* R32 = base + 6 * n * NR;
    float* D = base + 7 * n * NR;
    const unsigned all = (1u << NR) - 1;
    for (int o = 0;; ++o) {
        hipLaunchKernelGGL((k_mres<NR>), g, dim3(kWG), 0, h->stream, m64, N, B, X, R32, ost,
                           h->red, a.rel_tol, o == 0 ? 1 : 0);
        GIS_HIP(h, hipGetLastError());
        GIS_HIP(h, hipMemcpyAsync(ost_host, ost, sizeof(KState) * NR, hipMemcpyDeviceToHost,
                                  h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));
        unsigned skip = 0;
        for (int k = 0; k < NR; ++k) {
            if (ost_host[k].done) skip |= 1u << k;
            if (ost_host[k].status == kErrSync) h->red_rows = -1;   // refill the slots
        }
        if (skip == all) break;
        for (int k = 0; k < NR; ++k) {
            KState s{};
            // inner target: inner_rel_tol ||R||, but no further than half the outer
            // target (the last correction need not overshoot it)
            s.error = std::max(a.inner_rel_tol * ost_host[k].norm, 0.5 * ost_host[k].error);
            s.max_steps = a.max_inner <= 0 ? (long long)INT_MAX : a.max_inner;
            s.slot_fault = h->slot_fault;
            s.int_bound = a.int_bound;
            mb.host[k] = s;
        }
        GIS_HIP(h, hipMemcpyAsync(mb.st, mb.host, sizeof(KState) * NR, hipMemcpyHostToDevice,
                                  h->stream));
        GIS_HIP(h, hipMemsetAsync(D, 0, sizeof(float) * n * NR, h->stream));
        GIS_TRY((run_multi<NR, float>(h, method, m32, m32T, a.max_inner, D, R32, base, mb,
                                      skip)));
        hipLaunchKernelGGL((k_mcorr<NR>), g, dim3(kWG), 0, h->stream, N, X, D, mb.st, ost,
                           all & ~skip);
        GIS_HIP(h, hipGetLastError());
    }
    return kOK;
}

#define GIS_NR_DISPATCH(nrhs, NRC, ...)                     \
    switch (nrhs) {                                         \
        case 1: { constexpr int NRC = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int NRC = 2; __VA_ARGS__; } break; \
        case 3: { constexpr int NRC = 3; __VA_ARGS__; } break; \
        case 4: { constexpr int NRC = 4; __VA_ARGS__; } break; \
        case 5: { constexpr int NRC = 5; __VA_ARGS__; } break; \
        case 6: { constexpr int NRC = 6; __VA_ARGS__; } break; \
        case 7: { constexpr int NRC = 7; __VA_ARGS__; } break; \
        default: { constexpr int NRC = 8; __VA_ARGS__; } break; \
    }

}  // namespace

int device_solve_multi(gis_context* h, int method, int nrhs, bool transpose, const SolveArgs& a,
                       double* X, const double* B, int* iters, double* norm0, double* normf,
                       int* status) {
    if (nrhs < 1 || nrhs > kMaxRhs) return fail(h, kErrArg, "nrhs must be in 1..8");
    if (method != 0 && method != 1) return fail(h, kErrArg, "multi-RHS solves support CG and BiCG");
    GIS_TRY(ensure_valT(h));
    MultiBufs mb;
    GIS_TRY(multi_prepare(h, nrhs, a, sizeof(double) * (size_t)h->N * 6 * nrhs, mb));
    int st = kOK;
    GIS_NR_DISPATCH(nrhs, NRC, st = solve_multi_nr<NRC>(h, method, transpose, a, X, B, mb));
    if (st != kOK) return st;
    bool done = false;
    GIS_TRY(multi_read(h, nrhs, mb, &done));
    int worst = kOK;
    for (int k = 0; k < nrhs; ++k) {
        const KState& s = mb.host[k];
        int sk = s.status;
        if (sk == kOK && !(s.norm == s.norm)) sk = kNaN;
        if (iters) iters[k] = s.k;
        if (norm0) norm0[k] = s.norm0;
        if (normf) normf[k] = s.norm;
        if (status) status[k] = sk;
        worst = std::max(worst, sk);
    }
    return worst;
}

int device_solve_multi_mixed(gis_context* h, int method, int nrhs, const MultiMixedArgs& a,
                             double* X, const double* B, double* hist, int* outer, int* iters,
                             double* rel_res, int* status) {
    if (nrhs < 1 || nrhs > kMaxRhs) return fail(h, kErrArg, "nrhs must be in 1..8");
    if (method != 0 && method != 1) return fail(h, kErrArg, "multi-RHS solves support CG and BiCG");
    GIS_TRY(ensure_val32(h));
    if (method == 1) GIS_TRY(ensure_val32T(h));
    SolveArgs sa{};
    sa.hist_cap = 0;
    MultiBufs mb;
    // float workspace: 6 Krylov vectors, R32 and the correction D, per column
    GIS_TRY(multi_prepare(h, nrhs, sa, sizeof(float) * (size_t)h->N * 8 * nrhs, mb));
    const int cap = std::max(1, a.hist_cap);
    GIS_HIP(h, h->hist.ensure(sizeof(double) * (size_t)cap * nrhs));
    KState* ost = mb.st + kMaxRhs;
    KState* ost_host = mb.host + kMaxRhs;
    for (int k = 0; k < nrhs; ++k) {
        KState s{};
        s.slot_fault = h->slot_fault;
        s.max_steps = a.max_outer <= 0 ? (long long)INT_MAX : a.max_outer;
        s.hist = h->hist.as<double>() + (size_t)k * cap;
        s.hist_cap = cap;
        ost_host[k] = s;
    }
    GIS_HIP(h, hipMemcpyAsync(ost, ost_host, sizeof(KState) * nrhs, hipMemcpyHostToDevice,
                              h->stream));
    int st = kOK;
    GIS_NR_DISPATCH(nrhs, NRC,
                    st = solve_multi_mixed_nr<NRC>(h, method, a, X, B, mb, ost, ost_host));
    if (st != kOK) return st;
    GIS_HIP(h, hipMemcpyAsync(ost_host, ost, sizeof(KState) * nrhs, hipMemcpyDeviceToHost,
                              h->stream));
    if (hist && a.hist_cap > 0)
        GIS_HIP(h, hipMemcpyAsync(hist, h->hist.p, sizeof(double) * (size_t)cap * nrhs,
                                  hipMemcpyDeviceToHost, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    int worst = kOK;
    for (int k = 0; k < nrhs; ++k) {
        const KState& s = ost_host[k];
        if (outer) outer[k] = s.nhist;
        if (iters) iters[k] = (int)s.aux[0];
        if (rel_res) rel_res[k] = s.aux[3] > 0 ? s.norm / s.aux[3] : s.norm;
        if (status) status[k] = s.status;
        worst = std::max(worst, s.status);
    }
    return worst;
}

int device_sddmm_fold(gis_context* h, int nrhs, const double* Lam, const double* x,
                      double* d_diag, double* d_lower, double* d_upper) {
    hipLaunchKernelGGL(k_sddmm_fold, dim3((h->N + kWG - 1) / kWG), dim3(kWG), 0, h->stream, h->N,
                       h->M, nrhs, h->rowptr.as<int>(), h->col.as<int>(), h->perm.as<int>(), Lam,
                       x, d_diag, d_lower, d_upper);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

}  // namespace gis
