// Mixed-precision CG (no counterpart in the reference): the Krylov inner loop
// runs in FP32 (CSR values rounded to float, float SpMV products and row sums,
// float vector updates with float coefficients), every reduction accumulates
// in FP64, and FP64 is kept for the residual recompute r = b - A x, the
// correction accumulation x += (double) xc and the convergence norms.  Stops
// on the FP64 true relative residual ||b - A x||_2 / ||b||_2 <= rel_tol.
//
//   mode 0 (reliable updates): when the FP32 recursive residual falls below
//     inner_rel_tol times the last true residual (or below the target), the
//     correction is folded into x in FP64, the true residual replaces the FP32
//     one and the search direction continues with beta = rho_new / rho_old.
//   mode 1 (restarted refinement): same trigger, the inner CG restarts from p = r.
//
// Launches per inner iteration (each returns at once when it has nothing to do):
//   mx1: p_k = r + beta p_{k-1} on the fly, q = A32 p_k, deferred
//        xc += alpha_{k-1} p_{k-1}, (p_k, q) -> alpha_k
//   mx2: r -= alpha_k q, rho = (r, r), trigger test
//   mxu: (only after a trigger) x' = x + (double)(xc + alpha_k p_k) on the fly,
//        r64 = b - A x' in FP64, r = (float) r64, ||r64||, rho
// The deferred correction goes into xc in pairs: after a fold the first
// iteration leaves alpha_{k-1} p_{k-1} pending as well (F_PEND), the next one
// applies both, xc = fmaf(alpha_{k-1}, p_{k-1}, fmaf(alpha_{k-2}, p_{k-2}, xc)),
// the same float operations in the same order, so xc (and everything after it)
// keeps its bits while its read and write are saved every other iteration.
// The arithmetic is the one of oracle/gis_oracle_ext.c:or_mixed_cg, bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>

#include "gis_kernels.hpp"

namespace gis {

namespace {

enum : int {
    F_UPD = 1,       // FP64 update pending
    F_FOLDED = 2,    // xc already folded into x (next mx1 restarts xc at 0)
    F_RESTART = 4,   // next mx1 uses p = r
    F_XALT = 8,      // current x lives in the alternate buffer
    F_CONV = 16,     // converged
    F_PEND = 32,     // alpha_{k-1} p_{k-1} is pending too (aux[kAlphaPrevSlot])
};

// KState use: rho0 = rho, rho1 = rho_old, alpha, beta, norm (last true residual),
// norm0, error = target, aux[0] = inner_rel_tol, aux[1] = r0 (true norm at the
// last update), aux[3] = ||b||, max_steps = max_iter, aux[6] = mode,
// aux[kAlphaPrevSlot] = alpha_{k-1}.

// xc's pending terms with x_c = xc[c]: alpha_k p_k (pa = p_k), and with F_PEND
// alpha_{k-1} p_{k-1} (pb = p_{k-1}) applied first
__device__ __forceinline__ float pending_xc(bool pend, float af, float afp, float pa, float pb,
                                            float x_c) {
    return fmaf(af, pa, pend ? fmaf(afp, pb, x_c) : x_c);
}

// alpha_k from (p_k, q), and the flags for the next iteration: restart and
// fold are consumed; a pair of pending terms is opened after a plain
// iteration and closed by the next one
__device__ __forceinline__ void finish_alpha(KState* st, double pq) {
    const int f0 = st->flag;
    int f = f0 & ~(F_RESTART | F_FOLDED | F_PEND);
    if (!(f0 & (F_FOLDED | F_PEND))) f |= F_PEND;
    st->flag = f;
    st->aux[kAlphaPrevSlot] = st->alpha;
    st->alpha = st->rho0 / pq;
}

__device__ __forceinline__ void push_hist(KState* st, double v) {
    if (st->nhist < st->hist_cap) st->hist[st->nhist] = v;
    st->nhist = st->nhist + 1;
}

// set-up: r64 = b - A x (FP64), r32 = (float) r64, ||b||, ||r64||, rho = (r32, r32)
template <class P>
__global__ __launch_bounds__(kWG) void k_mx_setup(Mat<double> m, int N,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  float* __restrict__ r32, double rel_tol,
                                                  KState* st, RedWs red) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    double bv[4];
    float rf[4];
    load4(b, t.i0, N, bv);
    double leaf[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double r = bv[k] - sh_y[threadIdx.x * kLeaf + k];
        rf[k] = (float)r;
        if (t.i0 + k < N) {
            leaf[0] = fma(bv[k], bv[k], leaf[0]);
            leaf[1] = fma(r, r, leaf[1]);
            leaf[2] = fma((double)rf[k], (double)rf[k], leaf[2]);
        }
    }
    store4(r32, t.i0, N, rf);
    finish_tile_slots<3>(red, st, t.tile, leaf, shr, [&](const double (&v)[3]) {
        const double bnorm = sqrt(v[0]);
        const double rnorm = sqrt(v[1]);
        st->aux[3] = bnorm;
        st->error = rel_tol * bnorm;
        st->norm0 = rnorm;
        st->norm = rnorm;
        st->nhist = 0;
        push_hist(st, rnorm);
        st->rho0 = v[2];
        st->aux[1] = rnorm;
        st->k = 0;
        st->flag = F_RESTART | F_FOLDED;
        st->status = kOK;
        if (!(rnorm > st->error)) {
            st->done = 1;
            st->flag |= F_CONV;
        } else {
            st->done = 0;
        }
    });
}

template <class P, int C>
__global__ __launch_bounds__(kWG) void k_mx1(Mat<float> m, int N, const float* __restrict__ r,
                                             float* __restrict__ pb0, float* __restrict__ pb1,
                                             float* __restrict__ xc, float* __restrict__ q,
                                             KState* st, RedWs red) {
    if (st->done || (st->flag & F_UPD)) return;
    __shared__ typename P::template Shared<float> sh;
    __shared__ float sh_y[kTile];
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, false);
    const int k = st->k + 1;
    const int flag = st->flag;
    const bool first = (flag & F_RESTART) != 0;
    const bool folded = (flag & F_FOLDED) != 0;
    const bool pair = !folded && (flag & F_PEND) != 0;   // apply both pending terms
    const float bf = (float)st->beta;
    const float af = (float)st->alpha;
    const float afp = (float)st->aux[kAlphaPrevSlot];
    const float* p_old = (k & 1) ? pb0 : pb1;
    float* p_new = (k & 1) ? pb1 : pb0;
    double leaf[C][1];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        leaf[c][0] = 0.0;
        if (c >= sup.nt) continue;
        if (c > 0) __syncthreads();   // sh / sh_y reuse
        const TileRows t(N, sup.tile0 + c);
        if (first) {
            P::run(m, N, t.row0, t.nrows, [&](int j) { return r[j]; }, sh, sh_y);
        } else {
            P::run(m, N, t.row0, t.nrows, [&](int j) { return fmaf(bf, p_old[j], r[j]); }, sh,
                   sh_y);
        }
        float rv[4], pv[4], xv[4], pp[4], pn[4], qv[4];
        load4(r, t.i0, N, rv);
        if (!first || !folded) load4(p_old, t.i0, N, pv);
        if (pair) {
            load4nt(xc, t.i0, N, xv);
            load4(p_new, t.i0, N, pp);   // p_{k-2}, own rows, before p_k replaces it
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            pn[kk] = first ? rv[kk] : fmaf(bf, pv[kk], rv[kk]);
            qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
            xv[kk] = folded ? 0.0f : pair ? pending_xc(true, af, afp, pv[kk], pp[kk], xv[kk])
                                          : 0.0f;
            if (t.i0 + kk < N) leaf[c][0] = fma((double)pn[kk], (double)qv[kk], leaf[c][0]);
        }
        // non-temporal stores (and xc's load): measured at 256^3, 209.7 -> 205.2 us
        // per inner iteration against cached ones
        store4nt(p_new, t.i0, N, pn);
        store4nt(q, t.i0, N, qv);
        if (folded || pair) store4nt(xc, t.i0, N, xv);
    }
    SlotReduce<1, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
    tr.end(red, st, shr, [&](const double (&v)[1]) { finish_alpha(st, v[0]); });
}

// Split schedule, pass 1: the direction and the deferred correction of k_mx1,
// streamed (same float arithmetic per row).
template <int C>
__global__ __launch_bounds__(kWG) void k_mx_dir(int N, const float* __restrict__ r,
                                                float* __restrict__ pb0, float* __restrict__ pb1,
                                                float* __restrict__ xc, const KState* st) {
    if (st->done || (st->flag & F_UPD)) return;
    const SuperTile<C> sup(N, true);
    const int k = st->k + 1;
    const int flag = st->flag;
    const bool first = (flag & F_RESTART) != 0;
    const bool folded = (flag & F_FOLDED) != 0;
    const bool pair = !folded && (flag & F_PEND) != 0;
    const float bf = (float)st->beta;
    const float af = (float)st->alpha;
    const float afp = (float)st->aux[kAlphaPrevSlot];
    const float* p_old = (k & 1) ? pb0 : pb1;
    float* p_new = (k & 1) ? pb1 : pb0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (c >= sup.nt) continue;
        const TileRows t(N, sup.tile0 + c);
        float rv[4], pv[4] = {0.0f, 0.0f, 0.0f, 0.0f}, xv[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pn[4];
        float pp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        load4(r, t.i0, N, rv);
        if (!first || !folded) load4(p_old, t.i0, N, pv);
        if (pair) {
            load4(xc, t.i0, N, xv);
            load4(p_new, t.i0, N, pp);   // p_{k-2}, own rows, before p_k replaces it
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            pn[kk] = first ? rv[kk] : fmaf(bf, pv[kk], rv[kk]);
            xv[kk] = pair ? pending_xc(true, af, afp, pv[kk], pp[kk], xv[kk]) : 0.0f;
        }
        store4(p_new, t.i0, N, pn);
        if (folded || pair) store4(xc, t.i0, N, xv);
    }
}

// Split schedule, pass 2: q = A32 p_k with one gathered vector, (p_k, q) -> alpha_k.
template <class P, int C>
__global__ __launch_bounds__(kWG) void k_mx1s(Mat<float> m, int N, const float* __restrict__ pb0,
                                              const float* __restrict__ pb1,
                                              float* __restrict__ q, KState* st, RedWs red) {
    if (st->done || (st->flag & F_UPD)) return;
    __shared__ typename P::template Shared<float> sh;
    __shared__ float sh_y[kTile];
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, false);
    const float* p = ((st->k + 1) & 1) ? pb1 : pb0;
    double leaf[C][1];
    float qv[C][4];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        leaf[c][0] = 0.0;
        if (c >= sup.nt) continue;
        if (c > 0) __syncthreads();
        const TileRows t(N, sup.tile0 + c);
        P::run(m, N, t.row0, t.nrows, [&](int j) { return p[j]; }, sh, sh_y);
        float pv[4];
        load4(p, t.i0, N, pv);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            qv[c][kk] = sh_y[threadIdx.x * kLeaf + kk];
            if (t.i0 + kk < N) leaf[c][0] = fma((double)pv[kk], (double)qv[c][kk], leaf[c][0]);
        }
    }
    SlotReduce<1, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (c < sup.nt) store4nt(q, TileRows(N, sup.tile0 + c).i0, N, qv[c]);
    tr.end(red, st, shr, [&](const double (&v)[1]) { finish_alpha(st, v[0]); });
}

template <int C>
__global__ __launch_bounds__(kWG) void k_mx2(int N, float* __restrict__ r,
                                             const float* __restrict__ q, KState* st,
                                             RedWs red) {
    if (st->done || (st->flag & F_UPD)) return;
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, true);
    const float af = (float)st->alpha;
    double leaf[C][1];
    float rv[C][4];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        leaf[c][0] = 0.0;
        if (c >= sup.nt) continue;
        const TileRows t(N, sup.tile0 + c);
        float qv[4];
        load4(r, t.i0, N, rv[c]);
        // q is dead after this pass: a non-temporal load leaves the caches to
        // r, which the next k_mx1 gathers (35.9 against 37.3 us at 256^3)
        load4nt(q, t.i0, N, qv);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            rv[c][kk] = fmaf(-af, qv[kk], rv[c][kk]);
            if (t.i0 + kk < N) leaf[c][0] = fma((double)rv[c][kk], (double)rv[c][kk], leaf[c][0]);
        }
    }
    SlotReduce<1, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (c < sup.nt) store4(r, TileRows(N, sup.tile0 + c).i0, N, rv[c]);
    tr.end(red, st, shr, [&](const double (&v)[1]) {
        st->k = st->k + 1;
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        const double est = sqrt(v[0]);
        if (est != est) {
            st->status = kNaN;
            st->done = 1;
        } else if (est < st->aux[0] * st->aux[1] || !(est > st->error)) {
            st->flag = st->flag | F_UPD;
        } else {
            st->beta = st->rho0 / st->rho1;
            if (st->k >= st->max_steps) {
                st->status = kNotConverged;
                st->done = 1;
            }
        }
    });
}

// FP64 update; grid-strided over tiles so that the (usual) no-op launch is cheap.
// The grid is min(tiles, 1024): either one tile per workgroup, or a multiple
// of 64 workgroups, so that only workgroups with index 63 mod 64 ever collect a
// group of slots and every other workgroup runs through without waiting.
template <class P>
__global__ __launch_bounds__(kWG) void k_mxu(Mat<double> m, int N, const double* __restrict__ b,
                                             double* __restrict__ xb0, double* __restrict__ xb1,
                                             const float* __restrict__ pb0,
                                             const float* __restrict__ pb1,
                                             const float* __restrict__ xc,
                                             float* __restrict__ r32, KState* st, RedWs red) {
    if (st->done || !(st->flag & F_UPD)) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    const int k = st->k;
    const int flag = st->flag;
    const bool pend = (flag & F_PEND) != 0;
    const float af = (float)st->alpha;
    const float afp = (float)st->aux[kAlphaPrevSlot];
    const float* p = (k & 1) ? pb1 : pb0;
    const float* pq = (k & 1) ? pb0 : pb1;   // p_{k-1}, read only with F_PEND
    const double* x_old = (flag & F_XALT) ? xb1 : xb0;
    double* x_new = (flag & F_XALT) ? xb0 : xb1;
    auto xn = [&](int c) {
        return x_old[c] + (double)pending_xc(pend, af, afp, p[c], pend ? pq[c] : 0.0f, xc[c]);
    };
    const int ntiles = tiles_of(N);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        TileRows t(N, tile);
        P::run(m, N, t.row0, t.nrows, xn, sh, sh_y);
        double bv[4], xv[4];
        float rf[4];
        load4(b, t.i0, N, bv);
        double leaf[2] = {0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int i = t.i0 + kk;
            xv[kk] = i < N ? xn(i) : 0.0;
            const double r = bv[kk] - sh_y[threadIdx.x * kLeaf + kk];
            rf[kk] = (float)r;
            if (i < N) {
                leaf[0] = fma(r, r, leaf[0]);
                leaf[1] = fma((double)rf[kk], (double)rf[kk], leaf[1]);
            }
        }
        store4(x_new, t.i0, N, xv);
        store4(r32, t.i0, N, rf);
        finish_tile_slots<2>(red, st, t.tile, leaf, shr, [&](const double (&v)[2]) {
            const double rnorm = sqrt(v[0]);
            st->norm = rnorm;
            push_hist(st, rnorm);
            int f = (st->flag ^ F_XALT) & ~(F_UPD | F_PEND);
            f |= F_FOLDED;
            if (!(rnorm > st->error)) {
                f |= F_CONV;
                st->done = 1;
            } else {
                st->rho0 = v[1];
                st->aux[1] = rnorm;
                if (st->aux[6] != 0.0) f |= F_RESTART;
                else st->beta = st->rho0 / st->rho1;
                if (st->k >= st->max_steps) {
                    st->status = kNotConverged;
                    st->done = 1;
                }
            }
            st->flag = f;
        });
        __syncthreads();   // sh / sh_y reuse by the next tile
    }
}

// x_io = current x (+ pending correction when the loop stopped without converging)
__global__ __launch_bounds__(kWG) void k_mx_final(int N, double* __restrict__ xb0,
                                                  const double* __restrict__ xb1,
                                                  const float* __restrict__ pb0,
                                                  const float* __restrict__ pb1,
                                                  const float* __restrict__ xc,
                                                  const KState* st) {
    const int flag = st->flag;
    const int k = st->k;
    const bool conv = (flag & F_CONV) != 0;
    const bool folded = (flag & F_FOLDED) != 0 || k < 1;
    const bool pend = (flag & F_PEND) != 0;
    const float af = (float)st->alpha;
    const float afp = (float)st->aux[kAlphaPrevSlot];
    const float* p = (k & 1) ? pb1 : pb0;
    const float* pq = (k & 1) ? pb0 : pb1;
    const double* xcur = (flag & F_XALT) ? xb1 : xb0;
    StreamRows t(N);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int i = t.i0 + kk;
        if (i >= N) continue;
        double v = xcur[i];
        if (!conv)
            v = v + (double)(folded ? 0.0f
                                    : pending_xc(pend, af, afp, p[i], pend ? pq[i] : 0.0f, xc[i]));
        xb0[i] = v;
    }
}

}  // namespace

int device_solve_mixed(gis_context* h, const MixedArgs& a, double* x, const double* b,
                       int* hist_len, int* iters, double* rel_res) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_val32(h));
    SolveArgs sa{};
    sa.max_steps = a.max_iter;
    sa.hist_cap = a.hist_cap;
    GIS_TRY(prepare_solve(h, sa, 3));   // FP64: x alternate (+ room)
    const int N = h->N;
    const size_t n = (size_t)N;
    double* xb1 = h->vec.as<double>();
    // float workspace after the three FP64 vectors: r, p0, p1, q, xc
    float* f = reinterpret_cast<float*>(xb1 + 3 * n);
    GIS_HIP(h, h->vec.ensure(sizeof(double) * 3 * n + sizeof(float) * 5 * n));
    xb1 = h->vec.as<double>();
    f = reinterpret_cast<float*>(xb1 + 3 * n);
    float* r32 = f;
    float* pb0 = f + n;
    float* pb1 = f + 2 * n;
    float* q = f + 3 * n;
    float* xc = f + 4 * n;
    // mode and trigger ratio live in the state
    h->st_host->aux[0] = a.inner_rel_tol;
    h->st_host->aux[6] = (double)a.mode;
    GIS_HIP(h, hipMemcpyAsync(h->st, h->st_host, sizeof(KState), hipMemcpyHostToDevice,
                              h->stream));
    const Mat<double> m64 = mat64(h);
    const Mat<float> m32 = mat32(h);
    const dim3 g = tiles_grid(N);
    const dim3 gu((unsigned)std::min(tiles_of(N), 1024));
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        hipLaunchKernelGGL((k_mx_setup<P>), g, dim3(kWG), 0, h->stream, m64, N, b, x, r32,
                           a.rel_tol, h->st, h->red);
    });
    GIS_HIP(h, hipGetLastError());
    const long long max_iter = a.max_iter <= 0 ? (long long)INT_MAX : a.max_iter;
    // The FP64 update (k_mxu) follows every kGroup inner iterations instead of
    // every one: its launch is a no-op unless an inner iteration has raised
    // F_UPD, and the inner kernels launched after that raise return at once,
    // so the state sequence, and every result, is the same; the cap on inner
    // iterations is the device's (k >= max_steps), the host budget counts groups.
    constexpr int kGroup = 4;
    const long long max_groups = max_iter / kGroup + 1;
    GIS_SPMV_DISPATCH(h->sched.spmv32, P32, {
        GIS_SPMV_DISPATCH(h->sched.spmv64, P64, {
            const bool split = h->sched.fp32_direction != 0;
            GIS_COARSEN_DISPATCH(h->sched.coarsen32, CC, {
                const dim3 gc = super_grid(N, CC);
                GIS_TRY(run_blocks(h, max_groups, [&]() -> int {
                  for (int gi = 0; gi < kGroup; ++gi) {
                    if (split) {
                        GIS_TRY(timer_begin(h, 4));
                        // no reduction to amortise: one tile per workgroup (measured best)
                        hipLaunchKernelGGL((k_mx_dir<1>), g, dim3(kWG), 0, h->stream, N, r32,
                                           pb0, pb1, xc, h->st);
                        GIS_TRY(timer_end(h));
                        GIS_TRY(timer_begin(h, 0));
                        hipLaunchKernelGGL((k_mx1s<P32, CC>), gc, dim3(kWG), 0, h->stream, m32,
                                           N, pb0, pb1, q, h->st, h->red);
                    } else {
                        GIS_TRY(timer_begin(h, 0));
                        hipLaunchKernelGGL((k_mx1<P32, CC>), gc, dim3(kWG), 0, h->stream, m32, N,
                                           r32, pb0, pb1, xc, q, h->st, h->red);
                    }
                    GIS_TRY(timer_end(h));
                    GIS_TRY(timer_begin(h, 1));
                    hipLaunchKernelGGL((k_mx2<CC>), gc, dim3(kWG), 0, h->stream, N, r32, q,
                                       h->st, h->red);
                    GIS_TRY(timer_end(h));
                  }
                    GIS_TRY(timer_begin(h, 2));
                    hipLaunchKernelGGL((k_mxu<P64>), gu, dim3(kWG), 0, h->stream, m64, N, b, x,
                                       xb1, pb0, pb1, xc, r32, h->st, h->red);
                    GIS_TRY(timer_end(h));
                    return kOK;
                }));
            });
        });
    });
    hipLaunchKernelGGL(k_mx_final, g, dim3(kWG), 0, h->stream, N, x, xb1, pb0, pb1, xc, h->st);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(read_state(h));
    const KState& s = *h->st_host;
    if (hist_len) *hist_len = s.nhist;
    if (iters) *iters = s.k;
    if (rel_res) *rel_res = s.aux[3] > 0 ? s.norm / s.aux[3] : s.norm;
    return s.status;
}

}  // namespace gis
