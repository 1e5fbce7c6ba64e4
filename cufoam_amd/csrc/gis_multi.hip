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

// y[c][k] for the four 64-row slices of this wave, columns with bit k set in `active`
template <int NR, class XOp>
__device__ __forceinline__ void sell_multi(const Mat<double>& m, int N, int row0, XOp&& xop,
                                           unsigned active, double (&y)[4][NR]) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int k = 0; k < NR; ++k) y[c][k] = 0.0;
        const int s = s0 + c;
        if (s >= nslices) continue;
        const int a = m.soff[s];
        const int len = (m.soff[s + 1] - a) / kSlice;
        for (int j = 0; j < len; j += 4) {
            int cc[4];
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = a + (j + u) * kSlice + lane;
                if (j + u < len) {
                    cc[u] = __builtin_nontemporal_load(m.scol + idx);
                    v[u] = __builtin_nontemporal_load(m.sval + idx);
                } else {
                    cc[u] = -1;
                    v[u] = 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (cc[u] < 0) continue;
#pragma unroll
                for (int k = 0; k < NR; ++k)
                    if (active & (1u << k)) y[c][k] = y[c][k] + v[u] * xop(k, cc[u]);
            }
        }
    }
}

// column k of the wave's slices -> this thread's kLeaf consecutive rows
template <int NR>
__device__ __forceinline__ void to_leaf(const double (&y)[4][NR], int k, double* wave_lds,
                                        double (&o)[4]) {
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

// r_k = b_k - A x_k, optional copy, rho0 = (r, r), norm0, hist[0]
template <int NR>
__global__ __launch_bounds__(kWG) void k_msetup(Mat<double> m, int N, const double* __restrict__ B,
                                                const double* __restrict__ X,
                                                double* __restrict__ R, double* __restrict__ R2,
                                                KState* st, RedWs red, int done_rule) {
    __shared__ double sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    double y[4][NR];
    sell_multi<NR>(m, N, t.row0, [&](int k, int c) { return X[k * n + c]; }, (1u << NR) - 1, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double yl[4], bv[4], rv[4];
        to_leaf<NR>(y, k, sh_w[threadIdx.x >> 6], yl);
        load4(B + k * n, t.i0, N, bv);
        leaf[k] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rv[q] = bv[q] - yl[q];
            if (t.i0 + q < N) leaf[k] = fma(rv[q], rv[q], leaf[k]);
        }
        store4(R + k * n, t.i0, N, rv);
        if (R2) store4(R2 + k * n, t.i0, N, rv);
    }
    finish_tile<NR>(red, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
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
    });
}

// ---- directions (both methods), streamed ----------------------------------------
// p_k of iteration kk lives in buffer kk & 1.  CG: p = r + beta p_old and the
// deferred x += alpha_prev p_old (SolverSingleGPU.h:366-445).  BiCG: p = r +
// beta p_old, p* = r* + beta p* (SolverSingleGPU.h:568-675).  The SpMM that
// follows then gathers one vector per column instead of forming p at every
// gathered column from two.
template <int NR>
__global__ __launch_bounds__(kWG) void k_mdir(int N, const double* __restrict__ R,
                                              const double* __restrict__ RS,
                                              double* __restrict__ P0, double* __restrict__ P1,
                                              double* __restrict__ PS, double* __restrict__ X,
                                              const KState* st, int bicg) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    StreamRows t(N);
    const size_t n = (size_t)N;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (!(act & (1u << k))) continue;
        const int kk = st[k].k + 1;
        const double beta = st[k].beta;
        const double alpha_prev = st[k].alpha;
        const double* pold = ((kk & 1) ? P0 : P1) + k * n;
        double* pnew = ((kk & 1) ? P1 : P0) + k * n;
        double rv[4];
        load4(R + k * n, t.i0, N, rv);
        if (kk == 1) {
            store4(pnew, t.i0, N, rv);
            if (bicg) {
                double rsv[4];
                load4(RS + k * n, t.i0, N, rsv);
                store4(PS + k * n, t.i0, N, rsv);
            }
            continue;
        }
        double pv[4], pn[4];
        load4(pold, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q) pn[q] = fma(beta, pv[q], rv[q]);
        store4(pnew, t.i0, N, pn);
        if (bicg) {
            double rsv[4], psv[4];
            load4(RS + k * n, t.i0, N, rsv);
            load4(PS + k * n, t.i0, N, psv);
#pragma unroll
            for (int q = 0; q < 4; ++q) psv[q] = fma(beta, psv[q], rsv[q]);
            store4(PS + k * n, t.i0, N, psv);
        } else {
            double xv[4];
            load4(X + k * n, t.i0, N, xv);
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = fma(alpha_prev, pv[q], xv[q]);
            store4(X + k * n, t.i0, N, xv);
        }
    }
}

// ---- CG ---------------------------------------------------------------------
// after k_mdir: q = A p_k, (p_k, q) -> alpha_k
template <int NR>
__global__ __launch_bounds__(kWG) void k_mcg1(Mat<double> m, int N, const double* __restrict__ P0,
                                              const double* __restrict__ P1,
                                              double* __restrict__ Q, KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ double sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    // columns advance in lockstep until they stop: every active column is at
    // the same iteration, hence in the same buffer
    const double* P = ((st[__ffs(act) - 1].k + 1) & 1) ? P1 : P0;
    double y[4][NR];
    sell_multi<NR>(m, N, t.row0, [&](int k, int c) { return P[k * n + c]; }, act, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double yl[4];
        to_leaf<NR>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        double pv[4];
        load4(P + k * n, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t.i0 + q < N) leaf[k] = fma(pv[q], yl[q], leaf[k]);
        store4(Q + k * n, t.i0, N, yl);
    }
    finish_tile<NR>(red, t.tile, leaf, shr, [&](const double (&v)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (act & (1u << k)) st[k].alpha = st[k].rho0 / v[k];
    });
}

template <int NR>
__global__ __launch_bounds__(kWG) void k_mcg2(int N, double* __restrict__ R,
                                              const double* __restrict__ Q, KState* st,
                                              RedWs red) {
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
        const double alpha = st[k].alpha;
        double rv[4], qv[4];
        load4(R + k * n, t.i0, N, rv);
        load4(Q + k * n, t.i0, N, qv);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rv[q] = fma(-alpha, qv[q], rv[q]);
            if (t.i0 + q < N) leaf[k] = fma(rv[q], rv[q], leaf[k]);
        }
        store4(R + k * n, t.i0, N, rv);
    }
    finish_tile<NR>(red, t.tile, leaf, shr, [&](const double (&v)[NR]) {
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
    });
}

template <int NR>
__global__ __launch_bounds__(kWG) void k_mcg_flush(int N, double* __restrict__ X,
                                                   const double* __restrict__ P0,
                                                   const double* __restrict__ P1,
                                                   const KState* st) {
    StreamRows t(N);
    const size_t n = (size_t)N;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int it = st[k].k;
        if (it < 1) continue;
        const double alpha = st[k].alpha;
        const double* p = ((it & 1) ? P1 : P0) + k * n;
        double xv[4], pv[4];
        load4(X + k * n, t.i0, N, xv);
        load4(p, t.i0, N, pv);
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = fma(alpha, pv[q], xv[q]);
        store4(X + k * n, t.i0, N, xv);
    }
}

// ---- BiCG (the matrix m is "A", mt is its transpose) -------------------------
// after k_mdir: q = A p_k, (p*_k, q) -> xi = alpha_k, breakdown test
template <int NR>
__global__ __launch_bounds__(kWG) void k_mbi1(Mat<double> m, int N, const double* __restrict__ P0,
                                              const double* __restrict__ P1,
                                              const double* __restrict__ PS,
                                              double* __restrict__ Q, KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ double sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    const double* P = ((st[__ffs(act) - 1].k + 1) & 1) ? P1 : P0;
    double y[4][NR];
    sell_multi<NR>(m, N, t.row0, [&](int k, int c) { return P[k * n + c]; }, act, y);
    double leaf[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double yl[4];
        to_leaf<NR>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[k] = 0.0;
        if (!(act & (1u << k))) continue;
        double psv[4];
        load4(PS + k * n, t.i0, N, psv);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t.i0 + q < N) leaf[k] = fma(psv[q], yl[q], leaf[k]);
        store4(Q + k * n, t.i0, N, yl);
    }
    finish_tile<NR>(red, t.tile, leaf, shr, [&](const double (&v)[NR]) {
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
    });
}

template <int NR>
__global__ __launch_bounds__(kWG) void k_mbi2(Mat<double> mt, int N, double* __restrict__ R,
                                              double* __restrict__ RS,
                                              const double* __restrict__ P0,
                                              const double* __restrict__ P1,
                                              const double* __restrict__ PS,
                                              const double* __restrict__ Q,
                                              double* __restrict__ X, KState* st, RedWs red) {
    const unsigned act = active_mask<NR>(st);
    if (!act) return;
    __shared__ double sh_w[4][kTile / 4];
    __shared__ RedShared shr;
    TileRows t(N);
    const size_t n = (size_t)N;
    double y[4][NR];
    sell_multi<NR>(mt, N, t.row0, [&](int k, int c) { return PS[k * n + c]; }, act, y);
    double leaf[2 * NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double yl[4];
        to_leaf<NR>(y, k, sh_w[threadIdx.x >> 6], yl);
        leaf[2 * k] = 0.0;
        leaf[2 * k + 1] = 0.0;
        if (!(act & (1u << k))) continue;
        const double xi = st[k].alpha;
        const double* p = (((st[k].k + 1) & 1) ? P1 : P0) + k * n;
        double xv[4], pv[4], rv[4], qv[4], rsv[4];
        load4(X + k * n, t.i0, N, xv);
        load4(p, t.i0, N, pv);
        load4(R + k * n, t.i0, N, rv);
        load4(Q + k * n, t.i0, N, qv);
        load4(RS + k * n, t.i0, N, rsv);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xv[q] = fma(xi, pv[q], xv[q]);
            rv[q] = fma(-xi, qv[q], rv[q]);
            rsv[q] = fma(-xi, yl[q], rsv[q]);
            if (t.i0 + q < N) {
                leaf[2 * k] = fma(rv[q], rv[q], leaf[2 * k]);
                leaf[2 * k + 1] = fma(rv[q], rsv[q], leaf[2 * k + 1]);
            }
        }
        store4(X + k * n, t.i0, N, xv);
        store4(R + k * n, t.i0, N, rv);
        store4(RS + k * n, t.i0, N, rsv);
    }
    finish_tile<2 * NR>(red, t.tile, leaf, shr, [&](const double (&v)[2 * NR]) {
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
    });
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
            else d_upper[k * mm + (code - M - N)] = g;
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

int multi_prepare(gis_context* h, int nrhs, const SolveArgs& a, int nvec, MultiBufs& mb) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    GIS_HIP(h, h->vec.ensure(sizeof(double) * (size_t)h->N * nvec * nrhs));
    const int cap = std::max(1, a.hist_cap);
    GIS_HIP(h, h->hist.ensure(sizeof(double) * (size_t)cap * nrhs));
    GIS_HIP(h, h->mstate.ensure(sizeof(KState) * kMaxRhs));
    if (!h->mstate_host) {
        void* p = nullptr;
        GIS_HIP(h, hipHostMalloc(&p, sizeof(KState) * kMaxRhs, hipHostMallocDefault));
        h->mstate_host = static_cast<KState*>(p);
    }
    mb.st = h->mstate.as<KState>();
    mb.host = h->mstate_host;
    for (int k = 0; k < nrhs; ++k) {
        KState s{};
        s.error = a.error;
        s.max_steps = a.max_steps <= 0 ? (long long)INT_MAX : a.max_steps;
        s.int_bound = a.int_bound;
        s.hist = h->hist.as<double>() + (size_t)k * cap;
        s.hist_cap = cap;
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
    for (int k = 0; k < nrhs; ++k) d = d && mb.host[k].done;
    *all_done = d;
    return kOK;
}

template <int NR>
int solve_multi_nr(gis_context* h, int method, bool transpose, const SolveArgs& a, double* X,
                   const double* B, MultiBufs& mb) {
    const int N = h->N;
    const size_t n = (size_t)N;
    const Mat<double> mA = transpose ? mat64T(h) : mat64(h);
    const Mat<double> mAt = transpose ? mat64(h) : mat64T(h);
    const dim3 g = tiles_grid(N);
    double* base = h->vec.as<double>();
    double* R = base;
    double* P0 = base + 1 * n * NR;
    double* P1 = base + 2 * n * NR;
    double* Q = base + 3 * n * NR;
    double* RS = base + 4 * n * NR;
    double* PS = base + 5 * n * NR;
    const bool bicg = method == 1;
    hipLaunchKernelGGL((k_msetup<NR>), g, dim3(kWG), 0, h->stream, mA, N, B, X, R,
                       bicg ? RS : nullptr, mb.st, h->red, bicg ? 1 : 0);
    GIS_HIP(h, hipGetLastError());
    const long long ms = a.max_steps <= 0 ? (long long)INT_MAX : a.max_steps;
    long long launched = 0;
    int block = 4;
    for (;;) {
        for (int i = 0; i < block && launched < ms; ++i, ++launched) {
            GIS_TRY(timer_begin(h, 4));
            hipLaunchKernelGGL((k_mdir<NR>), g, dim3(kWG), 0, h->stream, N, R, RS, P0, P1, PS,
                               X, mb.st, bicg ? 1 : 0);
            GIS_TRY(timer_end(h));
            GIS_TRY(timer_begin(h, 0));
            if (bicg)
                hipLaunchKernelGGL((k_mbi1<NR>), g, dim3(kWG), 0, h->stream, mA, N, P0, P1, PS,
                                   Q, mb.st, h->red);
            else
                hipLaunchKernelGGL((k_mcg1<NR>), g, dim3(kWG), 0, h->stream, mA, N, P0, P1, Q,
                                   mb.st, h->red);
            GIS_TRY(timer_end(h));
            GIS_TRY(timer_begin(h, 1));
            if (bicg)
                hipLaunchKernelGGL((k_mbi2<NR>), g, dim3(kWG), 0, h->stream, mAt, N, R, RS, P0,
                                   P1, PS, Q, X, mb.st, h->red);
            else
                hipLaunchKernelGGL((k_mcg2<NR>), g, dim3(kWG), 0, h->stream, N, R, Q, mb.st,
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
        hipLaunchKernelGGL((k_mcg_flush<NR>), g, dim3(kWG), 0, h->stream, N, X, P0, P1, mb.st);
        GIS_HIP(h, hipGetLastError());
    }
    return kOK;
}

}  // namespace

int device_solve_multi(gis_context* h, int method, int nrhs, bool transpose, const SolveArgs& a,
                       double* X, const double* B, int* iters, double* norm0, double* normf,
                       int* status) {
    if (nrhs < 1 || nrhs > kMaxRhs) return fail(h, kErrArg, "nrhs must be in 1..8");
    if (method != 0 && method != 1) return fail(h, kErrArg, "multi-RHS solves support CG and BiCG");
    GIS_TRY(ensure_valT(h));
    MultiBufs mb;
    GIS_TRY(multi_prepare(h, nrhs, a, 6, mb));
    int st = kOK;
    switch (nrhs) {
        case 1: st = solve_multi_nr<1>(h, method, transpose, a, X, B, mb); break;
        case 2: st = solve_multi_nr<2>(h, method, transpose, a, X, B, mb); break;
        case 3: st = solve_multi_nr<3>(h, method, transpose, a, X, B, mb); break;
        case 4: st = solve_multi_nr<4>(h, method, transpose, a, X, B, mb); break;
        case 5: st = solve_multi_nr<5>(h, method, transpose, a, X, B, mb); break;
        case 6: st = solve_multi_nr<6>(h, method, transpose, a, X, B, mb); break;
        case 7: st = solve_multi_nr<7>(h, method, transpose, a, X, B, mb); break;
        default: st = solve_multi_nr<8>(h, method, transpose, a, X, B, mb); break;
    }
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

int device_sddmm_fold(gis_context* h, int nrhs, const double* Lam, const double* x,
                      double* d_diag, double* d_lower, double* d_upper) {
    hipLaunchKernelGGL(k_sddmm_fold, dim3((h->N + kWG - 1) / kWG), dim3(kWG), 0, h->stream, h->N,
                       h->M, nrhs, h->rowptr.as<int>(), h->col.as<int>(), h->perm.as<int>(), Lam,
                       x, d_diag, d_lower, d_upper);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

}  // namespace gis
