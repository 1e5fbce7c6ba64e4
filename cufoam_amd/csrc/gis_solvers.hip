// SpMV, reductions, NormFactor and CG on gfx950, FP64; host-side solver
// infrastructure (reduction workspace, device state, iteration blocks).
//
// Each method keeps its scalars (rho, alpha, ...) in a device KState and
// advances with a fixed sequence of launches per iteration; reductions finish
// inside the producing kernel (its slot collectors), so no iteration needs a
// host round-trip.  The host enqueues iterations in blocks and only
// reads the `done` flag between blocks; launches after convergence return at
// once, which preserves the reference's per-iteration stopping rule and
// therefore its iteration count.
#include <algorithm>
#include <climits>
#include <cmath>

#include "gis_kernels.hpp"

namespace gis {

// ---------------------------------------------------------------------------
// plain SpMV  y = A x  (SpMvKernel, kernels.h:206-237)
// ---------------------------------------------------------------------------
template <class P, typename T>
__global__ __launch_bounds__(kWG) void k_spmv(Mat<T> m, int N, const T* __restrict__ x,
                                              T* __restrict__ y) {
    __shared__ typename P::template Shared<T> sh;
    __shared__ T sh_y[kTile];
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    if constexpr (sizeof(T) == 8) {
        __syncthreads();
        store_tile_nt(y, sh_y, t.row0, N);
    } else {
        T o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = sh_y[threadIdx.x * kLeaf + k];
        store4nt(y, t.i0, N, o);   // 16 bytes of consecutive rows per thread: whole lines
    }
}

// ---------------------------------------------------------------------------
// reductions into KState::aux[slot]
// ---------------------------------------------------------------------------
// aux[slot] = dot(a, b)    (DotKernel, kernels.h:153-199)
__global__ __launch_bounds__(kWG) void k_dot(int N, const double* __restrict__ a,
                                             const double* __restrict__ b, KState* st,
                                             RedWs red, int slot) {
    __shared__ RedShared shr;
    StreamRows t(N);
    double av[4], bv[4];
    load4(a, t.i0, N, av);
    load4(b, t.i0, N, bv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = fma(av[k], bv[k], leaf[0]);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) { st->aux[slot] = v[0]; });
}

// aux[0] = sum(x) / N   (ReductionKernel kernels.h:100-145 + gis.cu:108)
__global__ __launch_bounds__(kWG) void k_mean(int N, const double* __restrict__ x, KState* st,
                                              RedWs red) {
    __shared__ RedShared shr;
    StreamRows t(N);
    double xv[4];
    load4(x, t.i0, N, xv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = leaf[0] + xv[k];
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) { st->aux[0] = v[0] / N; });
}

// NormFactor (gis/gis.cu:106-123):
//   aux[1] = ||A x - A x_bar||_2 + ||b - A x_bar||_2 + 1e-20, with x_bar = aux[0]
template <class P>
__global__ __launch_bounds__(kWG) void k_norm_factor(Mat<double> m, int N,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ x, KState* st,
                                                     RedWs red) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ double sh_yb[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const double x_bar = st->aux[0];
    P::run(m, N, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    P::run(m, N, t.row0, t.nrows, [&](int) { return x_bar; }, sh, sh_yb);
    double bv[4];
    load4(b, t.i0, N, bv);
    double leaf[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (t.i0 + k < N) {
            const int li = threadIdx.x * kLeaf + k;
            const double d1 = sh_y[li] - sh_yb[li];
            const double d2 = bv[k] - sh_yb[li];
            leaf[0] = fma(d1, d1, leaf[0]);
            leaf[1] = fma(d2, d2, leaf[1]);
        }
    }
    finish_tile_slots<2>(red, st, t.tile, leaf, shr, [&](const double (&v)[2]) {
        st->aux[1] = sqrt(v[0]) + sqrt(v[1]) + 1e-20;
    });
}

// ---------------------------------------------------------------------------
// CG  (SolverSingleGPU.h:366-445), two launches per iteration:
//   cg1: p_k = r + beta p_{k-1} (k > 1; formed on the fly, also at the gathered
//        columns), q = A p_k, x += alpha_{k-1} p_{k-1} (deferred), (p_k, q)
//        -> alpha_k = rho / (p_k, q)
//   cg2: r -= alpha_k q, rho = (r, r), norm, history, beta, stopping rule
// p alternates between two buffers (p_k in buffer k&1).
// ---------------------------------------------------------------------------
template <class P, int C>
__global__ __launch_bounds__(kWG) void k_cg1(Mat<double> m, int N, const double* __restrict__ r,
                                             double* __restrict__ pb0, double* __restrict__ pb1,
                                             double* __restrict__ x, double* __restrict__ q,
                                             KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, false);
    const int k = st->k + 1;
    const bool first = (k == 1);
    const double beta = st->beta;
    const double alpha_prev = st->alpha;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
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
            P::run(m, N, t.row0, t.nrows, [&](int j) { return fma(beta, p_old[j], r[j]); }, sh,
                   sh_y);
        }
        double rv[4], pv[4], xv[4], pn[4], qv[4];
        load4(r, t.i0, N, rv);
        if (!first) {
            load4(p_old, t.i0, N, pv);
            load4(x, t.i0, N, xv);
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            pn[kk] = first ? rv[kk] : fma(beta, pv[kk], rv[kk]);
            qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
            if (!first) xv[kk] = fma(alpha_prev, pv[kk], xv[kk]);
            if (t.i0 + kk < N) leaf[c][0] = fma(pn[kk], qv[kk], leaf[c][0]);
        }
        store4(p_new, t.i0, N, pn);
        if (!first) store4(x, t.i0, N, xv);
        __syncthreads();
        store_tile_nt(q, sh_y, t.row0, N);
    }
    SlotReduce<1, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
    tr.end(red, st, shr, [&](const double (&v)[1]) { st->alpha = st->rho0 / v[0]; });
}

// Split schedule, pass 1: p_k = r + beta p_{k-1} and the deferred x updates,
// the same fma per row as k_cg1, streamed.  The x updates go in pairs: an even
// iteration k leaves alpha_{k-1} p_{k-1} pending, and the odd iteration k + 1
// applies alpha_{k-1} p_{k-1}, then alpha_k p_k (p_{k-1} is still in the buffer
// p_{k+1} overwrites, read first row by row).  x takes the same fma in the same
// order, and a streamed read and write of x are saved every other iteration.
template <int C>
__global__ __launch_bounds__(kWG) void k_cg_dir(int N, const double* __restrict__ r,
                                                double* __restrict__ pb0,
                                                double* __restrict__ pb1,
                                                double* __restrict__ x, const KState* st) {
    if (st->done) return;
    const SuperTile<C> sup(N, true);
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    const double beta = st->beta;
    const double alpha_prev = st->alpha;                        // alpha_{k-1}
    const double alpha_prev2 = st->aux[kAlphaPrevSlot];         // alpha_{k-2}
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (c >= sup.nt) continue;
        const int row0 = (sup.tile0 + c) * kTile;
        double rv[4], pv[4], xv[4];   // rows as load_pairs maps them (no reduction here)
        load_pairs(r, row0, N, rv);
        if (k == 1) {
            store_pairs(p_new, row0, N, rv);
            continue;
        }
        load_pairs(p_old, row0, N, pv);
        if ((k & 1) == 0) {   // even: alpha_{k-1} p_{k-1} stays pending
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) pv[kk] = fma(beta, pv[kk], rv[kk]);
            store_pairs(p_new, row0, N, pv);
            continue;
        }
        double ppv[4];   // p_{k-2}, in the buffer p_k is about to take
        load_pairs(p_new, row0, N, ppv);
        load_pairs(x, row0, N, xv);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            xv[kk] = fma(alpha_prev2, ppv[kk], xv[kk]);
            xv[kk] = fma(alpha_prev, pv[kk], xv[kk]);
            pv[kk] = fma(beta, pv[kk], rv[kk]);
        }
        store_pairs(p_new, row0, N, pv);
        store_pairs(x, row0, N, xv);
    }
}

template <int C>
__global__ __launch_bounds__(kWG) void k_cg2(int N, double* __restrict__ r,
                                             const double* __restrict__ q, KState* st,
                                             RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, true);
    const double alpha = st->alpha;
    // rows in the pair layout (whole cache lines per access); the tile's sum of
    // squares in the canonical order by pair_tile_sumsq
    double tsum[C][1];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        tsum[c][0] = 0.0;
        if (c >= sup.nt) continue;
        const int row0 = (sup.tile0 + c) * kTile;
        double rv[4], qv[4];
        load_pairs(r, row0, N, rv);
        load_pairs(q, row0, N, qv);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) rv[kk] = fma(-alpha, qv[kk], rv[kk]);
        store_pairs_cached(r, row0, N, rv);   // read by k_cg_dir next
        tsum[c][0] = pair_tile_sumsq(rv, row0, N, shr);
    }
    SlotReduce<1, C> tr;
    tr.begin_sums(red, sup.tile0, sup.nt, tsum);
    tr.end(red, st, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = st->rho0 / st->rho1;
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (stop_after(st, !(norm > st->error))) st->done = 1;
    });
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// every published-partial slot to the empty pattern (SlotReduce)
__global__ void k_slot_fill(double* slot, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slot[i] = __longlong_as_double(static_cast<long long>(kSlotEmpty));
}

int fill_slots(gis_context* h, double* slot, size_t n) {
    if (n == 0) return kOK;
    hipLaunchKernelGGL(k_slot_fill, dim3((unsigned)((n + kWG - 1) / kWG)), dim3(kWG), 0,
                       h->stream, slot, (long long)n);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

// The level counts must describe exactly `rows` (the collectors wait for
// count[L] entries).  Level-0 and level-1 partials travel through slots;
// levels 2 and up (the top collector's global path) through plain partial
// arrays.  With base == nullptr only the size is computed.
bool reduction_layout(int rows, double* base, RedWs* out, size_t* n_doubles, size_t* n_slots) {
    RedWs r{};
    int count = tiles_of(rows);
    int L = 0;
    for (; L < kMaxLevels; ++L) {
        r.count[L] = count;
        if (count == 1) break;
        count = (count + kGroup - 1) / kGroup;
    }
    if (L == kMaxLevels) return false;
    r.nlevels = L + 1;
    for (int j = L + 1; j < kMaxLevels; ++j) r.count[j] = 1;
    size_t dbl = 0;
    for (int j = 2; j <= L; ++j) dbl += (size_t)r.count[j] * kMaxNV;
    const size_t nslot = ((size_t)r.count[0] + (size_t)r.count[1]) * kMaxNV;
    *n_doubles = dbl + nslot;
    *n_slots = nslot;
    if (!base) return true;
    r.slot = base;
    r.slot1 = base + (size_t)r.count[0] * kMaxNV;
    double* pp = base + nslot;
    for (int j = 0; j < kMaxLevels; ++j) {
        r.part[j] = nullptr;
        if (j >= 2 && j <= L) {
            r.part[j] = pp;
            pp += (size_t)r.count[j] * kMaxNV;
        }
    }
    *out = r;
    return true;
}

// Memory only grows; the slots are refilled whenever the shape changes.
int ensure_reduction(gis_context* h, int rows) {
    if (h->red_rows == rows && h->red_mem.p) return kOK;
    RedWs r{};
    size_t ndbl = 0, nslot = 0;
    if (!reduction_layout(rows, nullptr, &r, &ndbl, &nslot))
        return fail(h, kErrShape, "reduction depth exceeded");
    GIS_HIP(h, h->red_mem.ensure(ndbl * sizeof(double) + 256));
    reduction_layout(rows, h->red_mem.as<double>(), &r, &ndbl, &nslot);
    GIS_TRY(fill_slots(h, r.slot, nslot));
    h->red = r;
    h->red_rows = rows;
    return kOK;
}

// A fresh state for the one-shot reductions (dot, NormFactor): status GIS_OK,
// the handle's failure-path poll limit.
__global__ void k_state_reset(KState* st, int slot_fault) {
    if (threadIdx.x == 0) {
        KState s{};
        s.slot_fault = slot_fault;
        *st = s;
    }
}

int reset_state(gis_context* h, KState* st) {
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, h->stream, st, h->slot_fault);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int ensure_state(gis_context* h) {
    if (!h->st) {
        GIS_HIP(h, h->state_mem.ensure(sizeof(KState)));
        h->st = h->state_mem.as<KState>();
        GIS_HIP(h, hipMemsetAsync(h->st, 0, sizeof(KState), h->stream));
    }
    if (!h->st_host) {
        void* p = nullptr;
        GIS_HIP(h, hipHostMalloc(&p, sizeof(KState), hipHostMallocDefault));
        h->st_host = static_cast<KState*>(p);
    }
    return kOK;
}

int read_state(gis_context* h) {
    GIS_HIP(h, hipMemcpyAsync(h->st_host, h->st, sizeof(KState), hipMemcpyDeviceToHost,
                              h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    // a reduction whose wait ran out may have left slots written: rebuild the
    // workspace (and its empty pattern) before the next launch
    if (h->st_host->status == kErrSync) h->red_rows = -1;
    return timer_collect(h);
}

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------
int timer_begin(gis_context* h, int slot) {
    Timer& t = h->timer;
    t.open = false;
    if (!(t.mask & (1u << slot))) return kOK;
    // an event pair costs a few microseconds of launch gap: sample one launch in `period`
    if (t.seen[slot]++ % t.period != 0) return kOK;
    if ((size_t)(2 * t.used + 2) > t.ev.size()) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e;
            GIS_HIP(h, hipEventCreateWithFlags(&e, hipEventDefault));
            t.ev.push_back(e);
        }
        t.tag.push_back(0);
    }
    t.tag[t.used] = slot;
    GIS_HIP(h, hipEventRecord(t.ev[2 * t.used], h->stream));
    t.open = true;
    return kOK;
}

int timer_end(gis_context* h) {
    Timer& t = h->timer;
    if (!t.open) return kOK;
    GIS_HIP(h, hipEventRecord(t.ev[2 * t.used + 1], h->stream));
    t.used++;
    t.open = false;
    return kOK;
}

int timer_collect(gis_context* h) {
    Timer& t = h->timer;
    for (int i = 0; i < t.used; ++i) {
        GIS_HIP(h, hipEventSynchronize(t.ev[2 * i + 1]));
        float ms = 0.0f;
        GIS_HIP(h, hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]));
        t.total_ms[t.tag[i]] += ms;
        t.count[t.tag[i]] += 1;
    }
    t.used = 0;
    return kOK;
}

void timer_release(gis_context* h) {
    for (hipEvent_t e : h->timer.ev) (void)hipEventDestroy(e);
    h->timer.ev.clear();
    h->timer.tag.clear();
    h->timer.used = 0;
}

int launch_spmv(gis_context* h, const double* x, double* y) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    const Mat<double> m = mat64(h);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_spmv<P, double>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, m, h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int launch_spmv_t(gis_context* h, const double* x, double* y) {
    GIS_TRY(ensure_valT(h));
    const Mat<double> m = mat64T(h);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_spmv<P, double>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, m, h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int launch_spmv32(gis_context* h, const float* x, float* y) {
    GIS_TRY(ensure_val32(h));
    const Mat<float> m = mat32(h);
    GIS_SPMV_DISPATCH(h->sched.spmv32, P,
                      hipLaunchKernelGGL((k_spmv<P, float>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, m, h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int device_dot(gis_context* h, const double* a, const double* b, double* out) {
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    GIS_TRY(reset_state(h, h->st));
    hipLaunchKernelGGL(k_dot, tiles_grid(h->N), dim3(kWG), 0, h->stream, h->N, a, b, h->st,
                       h->red, 7);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(read_state(h));
    if (h->st_host->status == kErrSync) return fail(h, kErrSync, "dot: reduction did not complete");
    *out = h->st_host->aux[7];
    return kOK;
}

// x_bar into st->aux[0], then NormFactor into st->aux[1]; the workspace and
// the state are the caller's (the handle's, or the allocation-free scratch).
int launch_norm_factor(gis_context* h, const RedWs& red, KState* st, const double* d_b,
                       const double* d_x) {
    GIS_TRY(reset_state(h, st));
    hipLaunchKernelGGL(k_mean, tiles_grid(h->N), dim3(kWG), 0, h->stream, h->N, d_x, st, red);
    GIS_HIP(h, hipGetLastError());
    const Mat<double> m = mat64(h);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_norm_factor<P>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, m, h->N, d_b, d_x, st, red));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int device_norm_factor(gis_context* h, const double* d_b, const double* d_x, double* nf) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    GIS_TRY(launch_norm_factor(h, h->red, h->st, d_b, d_x));
    GIS_TRY(read_state(h));
    if (h->st_host->status == kErrSync)
        return fail(h, kErrSync, "NormFactor: reduction did not complete");
    *nf = h->st_host->aux[1];
    return kOK;
}

int prepare_solve(gis_context* h, const SolveArgs& a, int nvec) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    GIS_HIP(h, h->vec.ensure(sizeof(double) * (size_t)h->N * nvec));
    const int cap = std::max(1, a.hist_cap);
    GIS_HIP(h, h->hist.ensure(sizeof(double) * (size_t)cap));
    KState s{};
    s.error = a.error;
    s.max_steps = a.max_steps <= 0 ? (long long)INT_MAX : a.max_steps;
    s.omega = a.omega;
    s.int_bound = a.int_bound;
    s.rel_tol = a.rel_tol;
    s.min_steps = a.min_steps;
    s.slot_fault = h->slot_fault;
    s.detail = a.detail;
    s.tick = h->wall_tick;
    s.hist = h->hist.as<double>();
    s.hist_cap = cap;
    s.aux[4] = INFINITY;   // running min for FindDiagAbsMin
    *h->st_host = s;
    GIS_HIP(h, hipMemcpyAsync(h->st, h->st_host, sizeof(KState), hipMemcpyHostToDevice,
                              h->stream));
    return kOK;
}

int solve_cg(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(prepare_solve(h, a, 4));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* pb0 = r + N;
    double* pb1 = r + 2 * (size_t)N;
    double* q = r + 3 * (size_t)N;
    const dim3 g = tiles_grid(N);
    const Mat<double> m = mat64(h);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        hipLaunchKernelGGL((k_setup<P>), g, dim3(kWG), 0, h->stream, m, N, b, x, nullptr, r,
                           nullptr, h->st, h->red, 0);
        GIS_HIP(h, hipGetLastError());
        const bool split = h->sched.fp64_direction != 0;
        GIS_COARSEN_DISPATCH(h->sched.coarsen64, CC, {
            const dim3 gc = super_grid(N, CC);
            GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
                if (split) {
                    GIS_TRY(timer_begin(h, 4));
                    // no reduction to amortise: one tile per workgroup (measured best)
                    hipLaunchKernelGGL((k_cg_dir<1>), g, dim3(kWG), 0, h->stream, N, r, pb0,
                                       pb1, x, h->st);
                    GIS_TRY(timer_end(h));
                    GIS_TRY(timer_begin(h, 0));
                    hipLaunchKernelGGL((k_cg1s<P, CC>), gc, dim3(kWG), 0, h->stream, m, N, pb0,
                                       pb1, q, h->st, h->red);
                } else {
                    GIS_TRY(timer_begin(h, 0));
                    hipLaunchKernelGGL((k_cg1<P, CC>), gc, dim3(kWG), 0, h->stream, m, N, r, pb0,
                                       pb1, x, q, h->st, h->red);
                }
                GIS_TRY(timer_end(h));
                GIS_TRY(timer_begin(h, 1));
                hipLaunchKernelGGL((k_cg2<CC>), gc, dim3(kWG), 0, h->stream, N, r, q, h->st,
                                   h->red);
                GIS_TRY(timer_end(h));
                return kOK;
            }));
        });
    });
    hipLaunchKernelGGL(k_cg_flush, g, dim3(kWG), 0, h->stream, N, x, pb0, pb1, h->st,
                       h->sched.fp64_direction != 0 ? 1 : 0);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

}  // namespace gis
