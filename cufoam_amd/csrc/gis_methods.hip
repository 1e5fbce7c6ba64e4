// The seven non-CG methods of gis on gfx950, FP64, and the solve dispatcher.
//
// Reference loops (gis/SparseSystem/SolverSingleGPU.h):
//   BiCG :568-675   Richardson :762-819 (+OptimizePower :723-752,
//   FindDiagAbsMin :686-720)   MR :825-890   Jacobi :924-994
//   MRJacobi :1000-1073   JMR (PreconditionedMRJacobi) :1079-1155   MRS :1236-1303
//
// Every per-element formula matches the reference's kernels under nvcc's
// contraction (v = a + s*w -> fma(s, w, a), Hadamard v1*v2) and every
// reduction follows the fixed pairwise order of gis_device.hpp, so results
// are bit-identical to oracle/gis_oracle.c.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

#include "gis_kernels.hpp"

namespace gis {

int solve_cg(gis_context* h, const SolveArgs& a, double* x, const double* b);

namespace {

// ---------------------------------------------------------------------------
// per-row diagonal helpers (FindDiagInv / FindDiagInvSqrt, kernels.h:563-603):
// the LDU diagonal is the CSR diagonal, so they read diag[] directly.
//   mode 0: 1.0 / a_ii      mode 1: sqrt(1.0 / a_ii)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_diag_helper(int N, const double* __restrict__ diag,
                                                     double* __restrict__ out, int mode) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= N) return;
    const double inv = 1.0 / diag[i];
    out[i] = mode == 0 ? inv : sqrt(inv);
}

// min_i |a_ii| (FindDiag + MinAbsKernel, kernels.h:447-514): min is exact in any
// order; non-negative doubles order like their bit patterns, NaN bits sort above
// +inf and are ignored as fmin ignores them.
__global__ __launch_bounds__(kWG) void k_diag_absmin(int N, const double* __restrict__ diag,
                                                     unsigned long long* out) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    unsigned long long v = 0x7ff0000000000000ull;   // +inf
    if (i < N) v = (unsigned long long)__double_as_longlong(fabs(diag[i]));
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o = __shfl_xor(v, m, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) atomicMin(out, v);
}

// ---------------------------------------------------------------------------
// MR family: MR (z = r), MRJacobi (z = D^-1 r), MRS (z = 0 + invM r, invM =
// D^-1 as a diagonal CSR), JMR (z = S r, w = S A z, S = D^-1/2).
//   mr1: w = wscale o (A z), (r, w), (w, w) -> alpha = (r,w)/(w,w)
//   mr2: x = fma(alpha, z, x), r = fma(-alpha, w, r), (r, r) -> norm, stop
// ---------------------------------------------------------------------------
__device__ __forceinline__ double zform(const double* __restrict__ dvec, double r, int i,
                                        bool zero_add) {
    if (!dvec) return r;
    const double z = dvec[i] * r;
    return zero_add ? 0.0 + z : z;
}

template <class P>
__global__ __launch_bounds__(kWG) void k_mr1(Mat<double> m, int N, const double* __restrict__ r,
                                             const double* __restrict__ dvec,
                                             const double* __restrict__ wscale, int zero_add,
                                             double* __restrict__ w, KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const bool za = zero_add != 0;
    P::run(m, N, t.row0, t.nrows, [&](int c) { return zform(dvec, r[c], c, za); }, sh, sh_y);
    double rv[4], wv[4], sv[4];
    load4(r, t.i0, N, rv);
    if (wscale) load4(wscale, t.i0, N, sv);
    double leaf[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        wv[k] = sh_y[threadIdx.x * kLeaf + k];
        if (wscale) wv[k] = sv[k] * wv[k];
        if (t.i0 + k < N) {
            leaf[0] = fma(rv[k], wv[k], leaf[0]);
            leaf[1] = fma(wv[k], wv[k], leaf[1]);
        }
    }
    store4(w, t.i0, N, wv);
    finish_tile_slots<2>(red, st, t.tile, leaf, shr,
                   [&](const double (&v)[2]) { st->alpha = v[0] / v[1]; });
}

__global__ __launch_bounds__(kWG) void k_mr2(int N, double* __restrict__ r,
                                             const double* __restrict__ w,
                                             double* __restrict__ x,
                                             const double* __restrict__ dvec, int zero_add,
                                             KState* st, RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    StreamRows t(N);
    const double alpha = st->alpha;
    const bool za = zero_add != 0;
    double rv[4], wv[4], xv[4];
    load4(r, t.i0, N, rv);
    load4(w, t.i0, N, wv);
    load4(x, t.i0, N, xv);
    double dv[4] = {0.0, 0.0, 0.0, 0.0};
    if (dvec) load4(dvec, t.i0, N, dv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double z = rv[k];
        if (dvec) {
            z = dv[k] * rv[k];
            if (za) z = 0.0 + z;
        }
        xv[k] = fma(alpha, z, xv[k]);
        rv[k] = fma(-alpha, wv[k], rv[k]);
        if (t.i0 + k < N) leaf[0] = fma(rv[k], rv[k], leaf[0]);
    }
    store4(x, t.i0, N, xv);
    store4(r, t.i0, N, rv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (stop_after(st, !(norm > st->error))) st->done = 1;
    });
}

// ---------------------------------------------------------------------------
// Jacobi (residual form), one launch per iteration:
//   z = D^-1 r, w = A z, x = fma(1, z, x), r' = fma(-1, w, r), norm
// r alternates between two buffers (neighbours gather the old one).
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_jacobi(Mat<double> m, int N, double* __restrict__ rb0,
                                                double* __restrict__ rb1,
                                                const double* __restrict__ dinv,
                                                double* __restrict__ x, KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const int k = st->k + 1;
    const double* r_old = (k & 1) ? rb0 : rb1;
    double* r_new = (k & 1) ? rb1 : rb0;
    P::run(m, N, t.row0, t.nrows, [&](int c) { return dinv[c] * r_old[c]; }, sh, sh_y);
    double rv[4], dv[4], xv[4];
    load4(r_old, t.i0, N, rv);
    load4(dinv, t.i0, N, dv);
    load4(x, t.i0, N, xv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const double z = dv[kk] * rv[kk];
        xv[kk] = fma(1.0, z, xv[kk]);
        rv[kk] = fma(-1.0, sh_y[threadIdx.x * kLeaf + kk], rv[kk]);
        if (t.i0 + kk < N) leaf[0] = fma(rv[kk], rv[kk], leaf[0]);
    }
    store4(x, t.i0, N, xv);
    store4(r_new, t.i0, N, rv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (stop_after(st, !(norm > st->error))) st->done = 1;
    });
}

// ---------------------------------------------------------------------------
// Richardson, one launch per iteration (true residual every iteration):
//   x' = fma(omega, r, x) (formed on the fly at the gathered columns too),
//   r' = b - A x', norm.  x and r alternate between two buffers.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_richardson(Mat<double> m, int N,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ xb0,
                                                    double* __restrict__ xb1,
                                                    double* __restrict__ rb0,
                                                    double* __restrict__ rb1, KState* st,
                                                    RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const int k = st->k + 1;
    const double omega = st->omega;
    const double* x_old = (k & 1) ? xb0 : xb1;
    const double* r_old = (k & 1) ? rb0 : rb1;
    double* x_new = (k & 1) ? xb1 : xb0;
    double* r_new = (k & 1) ? rb1 : rb0;
    P::run(m, N, t.row0, t.nrows, [&](int c) { return fma(omega, r_old[c], x_old[c]); }, sh,
           sh_y);
    double rv[4], xv[4], bv[4];
    load4(r_old, t.i0, N, rv);
    load4(x_old, t.i0, N, xv);
    load4(b, t.i0, N, bv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        xv[kk] = fma(omega, rv[kk], xv[kk]);
        rv[kk] = bv[kk] - sh_y[threadIdx.x * kLeaf + kk];
        if (t.i0 + kk < N) leaf[0] = fma(rv[kk], rv[kk], leaf[0]);
    }
    store4(x_new, t.i0, N, xv);
    store4(r_new, t.i0, N, rv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (stop_after(st, !(norm > st->error))) st->done = 1;
    });
}

// OptimizePower step s: v = (s == 0 ? 1/N : aux[2] * t_in), t_out = A v,
// aux[2] = 1 / ||t_out||_2   (x1 = x2; x2 = A x1; x2 = (1/||x2||) x2)
template <class P>
__global__ __launch_bounds__(kWG) void k_power(Mat<double> m, int N,
                                               const double* __restrict__ t_in,
                                               double* __restrict__ t_out, KState* st,
                                               RedWs red) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const double alpha = st->aux[2];
    const double inv_n = 1.0 / N;
    if (t_in) {
        P::run(m, N, t.row0, t.nrows, [&](int c) { return alpha * t_in[c]; }, sh, sh_y);
    } else {
        P::run(m, N, t.row0, t.nrows, [&](int) { return inv_n; }, sh, sh_y);
    }
    double yv[4];
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        yv[k] = sh_y[threadIdx.x * kLeaf + k];
        if (t.i0 + k < N) leaf[0] = fma(yv[k], yv[k], leaf[0]);
    }
    store4(t_out, t.i0, N, yv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr,
                   [&](const double (&v)[1]) { st->aux[2] = 1 / sqrt(v[0]); });
}

// lambda_max = (x2, A x2); omega = 2 / (lambda_max + min|a_ii|)
template <class P>
__global__ __launch_bounds__(kWG) void k_power_final(Mat<double> m, int N,
                                                     const double* __restrict__ t_in,
                                                     KState* st, RedWs red) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const double alpha = st->aux[2];
    const double inv_n = 1.0 / N;
    auto v = [&](int c) { return t_in ? alpha * t_in[c] : inv_n; };
    P::run(m, N, t.row0, t.nrows, v, sh, sh_y);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = t.i0 + k;
        if (i < N) leaf[0] = fma(v(i), sh_y[threadIdx.x * kLeaf + k], leaf[0]);
    }
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&s)[1]) {
        const double lmax = s[0];
        const double m_abs = st->aux[4];   // written as bits by k_diag_absmin
        double lmin = DBL_MAX;
        if (lmin > m_abs) lmin = m_abs;
        st->aux[3] = lmax;
        st->aux[5] = lmin;
        st->omega = 2.0 / (lmax + lmin);
    });
}

// ---------------------------------------------------------------------------
// BiCG with A^T, two launches per iteration:
//   bi1: p_k = r + theta p_{k-1}, ps_k = rs + theta ps_{k-1} (k > 1),
//        q = A p_k, (ps_k, q) -> xi = rho / (ps_k, q), breakdown |xi| < bound
//   bi2: x += xi p_k, r -= xi q, rs -= xi A^T ps_k, (r, r) -> norm, history,
//        stop if norm < error; (r, rs) -> rho, theta
// p alternates between two buffers; rs and ps are only read row-locally
// where they are written, so they are updated in place.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_bicg1(Mat<double> m, int N, const double* __restrict__ r,
                                               const double* __restrict__ rs,
                                               double* __restrict__ pb0, double* __restrict__ pb1,
                                               double* __restrict__ ps, double* __restrict__ q,
                                               KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const int k = st->k + 1;
    const bool first = (k == 1);
    const double theta = st->beta;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    if (first) {
        P::run(m, N, t.row0, t.nrows, [&](int c) { return r[c]; }, sh, sh_y);
    } else {
        P::run(m, N, t.row0, t.nrows, [&](int c) { return fma(theta, p_old[c], r[c]); }, sh,
               sh_y);
    }
    double rv[4], rsv[4], pv[4], psv[4], pn[4], qv[4];
    load4(r, t.i0, N, rv);
    load4(rs, t.i0, N, rsv);
    if (!first) {
        load4(p_old, t.i0, N, pv);
        load4(ps, t.i0, N, psv);
    }
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        pn[kk] = first ? rv[kk] : fma(theta, pv[kk], rv[kk]);
        psv[kk] = first ? rsv[kk] : fma(theta, psv[kk], rsv[kk]);
        qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
        if (t.i0 + kk < N) leaf[0] = fma(psv[kk], qv[kk], leaf[0]);
    }
    store4(p_new, t.i0, N, pn);
    store4(ps, t.i0, N, psv);
    store4(q, t.i0, N, qv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double xi = st->rho0 / v[0];
        st->alpha = xi;
        if (xi > -st->int_bound && xi < st->int_bound) {
            st->status = kBreakdown;
            st->done = 1;
        }
    });
}

// Split schedule, pass 1: the two directions of k_bicg1 (same fma per row), streamed.
__global__ __launch_bounds__(kWG) void k_bicg_dir(int N, const double* __restrict__ r,
                                                  const double* __restrict__ rs,
                                                  double* __restrict__ pb0,
                                                  double* __restrict__ pb1,
                                                  double* __restrict__ ps, const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    double rv[4], rsv[4];
    load4(r, t.i0, N, rv);
    load4(rs, t.i0, N, rsv);
    if (k == 1) {
        store4(p_new, t.i0, N, rv);
        store4(ps, t.i0, N, rsv);
        return;
    }
    const double theta = st->beta;
    double pv[4], psv[4];
    load4(p_old, t.i0, N, pv);
    load4(ps, t.i0, N, psv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        pv[kk] = fma(theta, pv[kk], rv[kk]);
        psv[kk] = fma(theta, psv[kk], rsv[kk]);
    }
    store4(p_new, t.i0, N, pv);
    store4(ps, t.i0, N, psv);
}

// Split schedule, pass 2: q = A p_k with one gathered vector, (p*_k, q) -> xi.
template <class P>
__global__ __launch_bounds__(kWG) void k_bicg1s(Mat<double> m, int N,
                                                const double* __restrict__ pb0,
                                                const double* __restrict__ pb1,
                                                const double* __restrict__ ps,
                                                double* __restrict__ q, KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const double* p = ((st->k + 1) & 1) ? pb1 : pb0;
    P::run(m, N, t.row0, t.nrows, [&](int c) { return p[c]; }, sh, sh_y);
    double psv[4], qv[4];
    load4(ps, t.i0, N, psv);
    double leaf[1][1] = {{0.0}};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
        if (t.i0 + kk < N) leaf[0][0] = fma(psv[kk], qv[kk], leaf[0][0]);
    }
    SlotReduce<1, 1> tr;
    tr.begin(red, t.tile, 1, leaf, shr);
    store4(q, t.i0, N, qv);
    tr.end(red, st, shr, [&](const double (&v)[1]) {
        const double xi = st->rho0 / v[0];
        st->alpha = xi;
        if (xi > -st->int_bound && xi < st->int_bound) {
            st->status = kBreakdown;
            st->done = 1;
        }
    });
}

// RHO 0: rho = (r, rs); 1: rho = (D^-1 r, rs) from dinv (BiCG preconditioned by
// M = D); 2: rho is left to a later kernel (k_pbicg_rz, after M^-1 r and M^-T rs).
template <class P, int RHO>
__global__ __launch_bounds__(kWG) void k_bicg2(Mat<double> mt, int N, double* __restrict__ r,
                                               double* __restrict__ rs,
                                               const double* __restrict__ pb0,
                                               const double* __restrict__ pb1,
                                               const double* __restrict__ ps,
                                               const double* __restrict__ q,
                                               const double* __restrict__ dinv,
                                               double* __restrict__ x, KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const int k = st->k + 1;
    const double xi = st->alpha;
    const double* p = (k & 1) ? pb1 : pb0;
    P::run(mt, N, t.row0, t.nrows, [&](int c) { return ps[c]; }, sh, sh_y);
    double xv[4], pv[4], rv[4], qv[4], rsv[4], dv[4];
    load4(x, t.i0, N, xv);
    load4(p, t.i0, N, pv);
    load4(r, t.i0, N, rv);
    load4(q, t.i0, N, qv);
    load4(rs, t.i0, N, rsv);
    if (RHO == 1) load4(dinv, t.i0, N, dv);
    double leaf[2] = {0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        xv[kk] = fma(xi, pv[kk], xv[kk]);
        rv[kk] = fma(-xi, qv[kk], rv[kk]);
        rsv[kk] = fma(-xi, sh_y[threadIdx.x * kLeaf + kk], rsv[kk]);
        if (t.i0 + kk < N) {
            leaf[0] = fma(rv[kk], rv[kk], leaf[0]);
            if (RHO < 2) leaf[1] = fma(RHO == 1 ? dv[kk] * rv[kk] : rv[kk], rsv[kk], leaf[1]);
        }
    }
    store4(x, t.i0, N, xv);
    store4(r, t.i0, N, rv);
    store4(rs, t.i0, N, rsv);
    finish_tile_slots<2>(red, st, t.tile, leaf, shr, [&](const double (&v)[2]) {
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (stop_after(st, norm < st->error)) {
            st->done = 1;
        } else if (RHO < 2) {
            st->rho1 = st->rho0;
            st->rho0 = v[1];
            st->beta = st->rho0 / st->rho1;
        }
    });
}

// ---------------------------------------------------------------------------
// Preconditioned BiCG with M = D (split schedule): z = D^-1 r, z* = D^-1 r*,
//   pdir: p_k = z (k = 1) or fma(theta, p_{k-1}, z), ps_k likewise from z*
//   k_bicg1s: q = A p_k, (ps_k, q) -> xi, breakdown
//   k_bicg2<P, true>: x, r, rs as BiCG; (r, r) -> norm; (D^-1 r, rs) -> rho
// rho0 = (D^-1 r0, r0) after the set-up (rs0 = r0).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWG) void k_pbicg_rho0(int N, const double* __restrict__ dinv,
                                                    const double* __restrict__ r, KState* st,
                                                    RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
    double dv[4], rv[4];
    load4(dinv, t.i0, N, dv);
    load4(r, t.i0, N, rv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        if (t.i0 + kk < N) leaf[0] = fma(dv[kk] * rv[kk], rv[kk], leaf[0]);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) { st->rho0 = v[0]; });
}

__global__ __launch_bounds__(kWG) void k_pbicg_dir(int N, const double* __restrict__ dinv,
                                                   const double* __restrict__ r,
                                                   const double* __restrict__ rs,
                                                   double* __restrict__ pb0,
                                                   double* __restrict__ pb1,
                                                   double* __restrict__ ps, const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    double dv[4], zv[4], zsv[4];
    load4(dinv, t.i0, N, dv);
    load4(r, t.i0, N, zv);
    load4(rs, t.i0, N, zsv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        zv[kk] = dv[kk] * zv[kk];
        zsv[kk] = dv[kk] * zsv[kk];
    }
    if (k == 1) {
        store4(p_new, t.i0, N, zv);
        store4(ps, t.i0, N, zsv);
        return;
    }
    const double theta = st->beta;
    double pv[4], psv[4];
    load4(p_old, t.i0, N, pv);
    load4(ps, t.i0, N, psv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        pv[kk] = fma(theta, pv[kk], zv[kk]);
        psv[kk] = fma(theta, psv[kk], zsv[kk]);
    }
    store4(p_new, t.i0, N, pv);
    store4(ps, t.i0, N, psv);
}

// Preconditioned BiCG with a general M (DILU): the directions from z = M^-1 r
// and zs = M^-T rs, formed by the sweeps; rho = (z, rs) by k_pbicg_rz.
__global__ __launch_bounds__(kWG) void k_pbicg_dirz(int N, const double* __restrict__ z,
                                                    const double* __restrict__ zs,
                                                    double* __restrict__ pb0,
                                                    double* __restrict__ pb1,
                                                    double* __restrict__ ps, const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    double zv[4], zsv[4];
    load4(z, t.i0, N, zv);
    load4(zs, t.i0, N, zsv);
    if (k == 1) {
        store4(p_new, t.i0, N, zv);
        store4(ps, t.i0, N, zsv);
        return;
    }
    const double theta = st->beta;
    double pv[4], psv[4];
    load4(p_old, t.i0, N, pv);
    load4(ps, t.i0, N, psv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        pv[kk] = fma(theta, pv[kk], zv[kk]);
        psv[kk] = fma(theta, psv[kk], zsv[kk]);
    }
    store4(p_new, t.i0, N, pv);
    store4(ps, t.i0, N, psv);
}

// rho = (z, rs); FIRST: the set-up's rho, else also theta = rho / rho_prev
template <int FIRST>
__global__ __launch_bounds__(kWG) void k_pbicg_rz(int N, const double* __restrict__ z,
                                                  const double* __restrict__ rs, KState* st,
                                                  RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
    double zv[4], rsv[4];
    load4(z, t.i0, N, zv);
    load4(rs, t.i0, N, rsv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        if (t.i0 + kk < N) leaf[0] = fma(zv[kk], rsv[kk], leaf[0]);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        if (!FIRST) st->rho1 = st->rho0;
        st->rho0 = v[0];
        if (!FIRST) st->beta = st->rho0 / st->rho1;
    });
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
int launch_setup(gis_context* h, const Mat<double>& m, const double* b, const double* x,
                 const double* scale, double* r, double* r2, int done_rule) {
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_setup<P>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, m, h->N, b, x, scale, r, r2, h->st, h->red,
                                         done_rule));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int diag_helper(gis_context* h, int mode) {
    GIS_HIP(h, h->dinv.ensure(sizeof(double) * (size_t)h->N));
    hipLaunchKernelGGL(k_diag_helper, dim3((h->N + kWG - 1) / kWG), dim3(kWG), 0, h->stream,
                       h->N, h->diag.as<double>(), h->dinv.as<double>(), mode);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

// MR, MRJacobi, MRS, JMR
int solve_mr_family(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(prepare_solve(h, a, 2));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* w = r + N;
    const double* dvec = nullptr;
    const double* wscale = nullptr;
    const double* rscale = nullptr;
    int zero_add = 0;
    if (a.method == 4 || a.method == 6) {   // MRJacobi, MRS: D^-1
        GIS_TRY(diag_helper(h, 0));
        dvec = h->dinv.as<double>();
        zero_add = a.method == 6;
    } else if (a.method == 5) {             // JMR: S = sqrt(D^-1)
        GIS_TRY(diag_helper(h, 1));
        dvec = wscale = rscale = h->dinv.as<double>();
    }
    const Mat<double> m = mat64(h);
    GIS_TRY(launch_setup(h, m, b, x, rscale, r, nullptr, 0));
    const dim3 g = tiles_grid(N);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            hipLaunchKernelGGL((k_mr1<P>), g, dim3(kWG), 0, h->stream, m, N, r, dvec, wscale,
                               zero_add, w, h->st, h->red);
            hipLaunchKernelGGL(k_mr2, g, dim3(kWG), 0, h->stream, N, r, w, x, dvec, zero_add,
                               h->st, h->red);
            return kOK;
        }));
    });
    return kOK;
}

int solve_jacobi(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(prepare_solve(h, a, 2));
    const int N = h->N;
    double* rb0 = h->vec.as<double>();
    double* rb1 = rb0 + N;
    GIS_TRY(diag_helper(h, 0));
    const Mat<double> m = mat64(h);
    GIS_TRY(launch_setup(h, m, b, x, nullptr, rb0, nullptr, 0));
    const dim3 g = tiles_grid(N);
    const double* dinv = h->dinv.as<double>();
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            hipLaunchKernelGGL((k_jacobi<P>), g, dim3(kWG), 0, h->stream, m, N, rb0, rb1, dinv,
                               x, h->st, h->red);
            return kOK;
        }));
    });
    return kOK;
}

int solve_richardson(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(prepare_solve(h, a, 5));
    const int N = h->N;
    double* rb0 = h->vec.as<double>();
    double* rb1 = rb0 + N;
    double* xb1 = rb0 + 2 * (size_t)N;
    double* t0 = rb0 + 3 * (size_t)N;
    double* t1 = rb0 + 4 * (size_t)N;
    const Mat<double> m = mat64(h);
    const dim3 g = tiles_grid(N);
    if (-1e-15 < a.omega && a.omega < 1e-15) {
        // OptimizePower: aux[4] collects min|a_ii| (prepare_solve starts it at +inf)
        hipLaunchKernelGGL(k_diag_absmin, dim3((N + kWG - 1) / kWG), dim3(kWG), 0, h->stream,
                           N, h->diag.as<double>(),
                           reinterpret_cast<unsigned long long*>(&h->st->aux[4]));
        GIS_HIP(h, hipGetLastError());
        const double* tin = nullptr;
        GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
            for (int s = 0; s < a.power_steps; ++s) {
                double* tout = (s & 1) ? t1 : t0;
                hipLaunchKernelGGL((k_power<P>), g, dim3(kWG), 0, h->stream, m, N, tin, tout,
                                   h->st, h->red);
                tin = tout;
            }
            hipLaunchKernelGGL((k_power_final<P>), g, dim3(kWG), 0, h->stream, m, N, tin,
                               h->st, h->red);
        });
        GIS_HIP(h, hipGetLastError());
    }
    GIS_TRY(launch_setup(h, m, b, x, nullptr, rb0, nullptr, 0));
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            hipLaunchKernelGGL((k_richardson<P>), g, dim3(kWG), 0, h->stream, m, N, b, x, xb1,
                               rb0, rb1, h->st, h->red);
            return kOK;
        }));
    });
    if (h->st_host->k & 1)
        GIS_HIP(h, hipMemcpyAsync(x, xb1, sizeof(double) * (size_t)N, hipMemcpyDeviceToDevice,
                                  h->stream));
    return kOK;
}

int solve_bicg(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(ensure_valT(h));
    GIS_TRY(prepare_solve(h, a, 6));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* rs = r + N;
    double* pb0 = r + 2 * (size_t)N;
    double* pb1 = r + 3 * (size_t)N;
    double* ps = r + 4 * (size_t)N;
    double* q = r + 5 * (size_t)N;
    const Mat<double> m = mat64(h);
    const Mat<double> mt = mat64T(h);
    GIS_TRY(launch_setup(h, m, b, x, nullptr, r, rs, 1));
    const dim3 g = tiles_grid(N);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            if (h->sched.fp64_direction) {
                hipLaunchKernelGGL(k_bicg_dir, g, dim3(kWG), 0, h->stream, N, r, rs, pb0, pb1, ps,
                                   h->st);
                hipLaunchKernelGGL((k_bicg1s<P>), g, dim3(kWG), 0, h->stream, m, N, pb0, pb1, ps,
                                   q, h->st, h->red);
            } else {
                hipLaunchKernelGGL((k_bicg1<P>), g, dim3(kWG), 0, h->stream, m, N, r, rs, pb0,
                                   pb1, ps, q, h->st, h->red);
            }
            hipLaunchKernelGGL((k_bicg2<P, 0>), g, dim3(kWG), 0, h->stream, mt, N, r, rs,
                               pb0, pb1, ps, q, nullptr, x, h->st, h->red);
            return kOK;
        }));
    });
    return kOK;
}

// BiCG preconditioned by M = D (no counterpart in the reference): the BiCG
// loop, breakdown test, stopping rule and history, with z = D^-1 r in the
// directions and in rho.  The split schedule only (three launches per iteration).
int solve_pbicg_diag(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(ensure_valT(h));
    GIS_TRY(prepare_solve(h, a, 6));
    GIS_TRY(diag_helper(h, 0));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* rs = r + N;
    double* pb0 = r + 2 * (size_t)N;
    double* pb1 = r + 3 * (size_t)N;
    double* ps = r + 4 * (size_t)N;
    double* q = r + 5 * (size_t)N;
    const double* dinv = h->dinv.as<double>();
    const Mat<double> m = mat64(h);
    const Mat<double> mt = mat64T(h);
    GIS_TRY(launch_setup(h, m, b, x, nullptr, r, rs, 1));
    const dim3 g = tiles_grid(N);
    hipLaunchKernelGGL(k_pbicg_rho0, g, dim3(kWG), 0, h->stream, N, dinv, r, h->st, h->red);
    GIS_HIP(h, hipGetLastError());
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            hipLaunchKernelGGL(k_pbicg_dir, g, dim3(kWG), 0, h->stream, N, dinv, r, rs, pb0,
                               pb1, ps, h->st);
            hipLaunchKernelGGL((k_bicg1s<P>), g, dim3(kWG), 0, h->stream, m, N, pb0, pb1, ps,
                               q, h->st, h->red);
            hipLaunchKernelGGL((k_bicg2<P, 1>), g, dim3(kWG), 0, h->stream, mt, N, r, rs,
                               pb0, pb1, ps, q, dinv, x, h->st, h->red);
            return kOK;
        }));
    });
    return kOK;
}

// BiCG preconditioned by OpenFOAM's DILU (no counterpart in the reference):
// z = M^-1 r and zs = M^-T rs by the level-scheduled sweeps of gis_amg.hip (the
// forward sweep reads lower, the backward one upper; M^-T exchanges them).
int solve_pbicg_dilu(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(ensure_valT(h));
    GIS_TRY(ilu_prepare(h, 5));
    GIS_TRY(prepare_solve(h, a, 8));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* rs = r + N;
    double* pb0 = r + 2 * (size_t)N;
    double* pb1 = r + 3 * (size_t)N;
    double* ps = r + 4 * (size_t)N;
    double* q = r + 5 * (size_t)N;
    double* z = r + 6 * (size_t)N;
    double* zs = r + 7 * (size_t)N;
    const double* lo = h->lower.as<double>();
    const double* up = h->upper.as<double>();
    const Mat<double> m = mat64(h);
    const Mat<double> mt = mat64T(h);
    GIS_TRY(launch_setup(h, m, b, x, nullptr, r, rs, 1));
    const dim3 g = tiles_grid(N);
    GIS_TRY(ilu_apply(h, r, z, lo, up));
    GIS_TRY(ilu_apply(h, rs, zs, up, lo));
    hipLaunchKernelGGL(k_pbicg_rz<1>, g, dim3(kWG), 0, h->stream, N, z, rs, h->st, h->red);
    GIS_HIP(h, hipGetLastError());
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
            hipLaunchKernelGGL(k_pbicg_dirz, g, dim3(kWG), 0, h->stream, N, z, zs, pb0, pb1, ps,
                               h->st);
            hipLaunchKernelGGL((k_bicg1s<P>), g, dim3(kWG), 0, h->stream, m, N, pb0, pb1, ps,
                               q, h->st, h->red);
            hipLaunchKernelGGL((k_bicg2<P, 2>), g, dim3(kWG), 0, h->stream, mt, N, r, rs,
                               pb0, pb1, ps, q, nullptr, x, h->st, h->red);
            GIS_TRY(ilu_apply(h, r, z, lo, up));
            GIS_TRY(ilu_apply(h, rs, zs, up, lo));
            hipLaunchKernelGGL(k_pbicg_rz<0>, g, dim3(kWG), 0, h->stream, N, z, rs, h->st,
                               h->red);
            return kOK;
        }, true));
    });
    return kOK;
}

}  // namespace

int device_solve(gis_context* h, const SolveArgs& a, double* d_x, const double* d_b,
                 int* hist_len, double* norm0, double* norm_final) {
    int st;
    switch (a.method) {
        case 0: st = solve_cg(h, a, d_x, d_b); break;
        case 1: st = solve_bicg(h, a, d_x, d_b); break;
        case 2: st = solve_jacobi(h, a, d_x, d_b); break;
        case 3:
        case 4:
        case 5:
        case 6: st = solve_mr_family(h, a, d_x, d_b); break;
        case 7: st = solve_richardson(h, a, d_x, d_b); break;
        case kMethodPcg0 + 0:
        case kMethodPcg0 + 1:
        case kMethodPcg0 + 2:
        case kMethodPcg0 + 3:
        case kMethodPcg0 + 4: st = solve_pcg(h, a, a.method - kMethodPcg0, d_x, d_b); break;
        case kMethodPbicgDiag: st = solve_pbicg_diag(h, a, d_x, d_b); break;
        case kMethodPbicgDilu: st = solve_pbicg_dilu(h, a, d_x, d_b); break;
        default: return fail(h, kErrArg, "unknown solver method");
    }
    if (st != kOK) return st;
    GIS_TRY(read_state(h));
    const KState& s = *h->st_host;
    if (hist_len) *hist_len = s.k + 1;
    if (norm0) *norm0 = s.norm0;
    if (norm_final) *norm_final = s.norm;
    if (s.status != kOK) return s.status;
    if (!(s.norm == s.norm)) return kNaN;
    return kOK;
}

}  // namespace gis
