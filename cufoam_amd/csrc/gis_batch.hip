// Batched CG over many independent systems of uneven size (no counterpart in
// the reference, which solves one matrix per call).
//
// Layout: the systems are stacked block-diagonally; system s owns rows
// [off[s], off[s] + n[s]) and is padded with empty rows (zero coefficients,
// zero vectors) to a whole number of kTile-row tiles, so no tile straddles two
// systems and the tiles of one system form its own reduction tree.  Padding
// rows only add zero leaves, which leave the pairwise tree unchanged, so every
// system reproduces its single-system CG (SolverSingleGPU.h:366-445) bit for
// bit, including its own iteration count and stopping rule.
//
// One launch per CG stage serves all systems (a workgroup finds its system
// from its tile); converged systems' workgroups return at once.  The stages of
// a block of iterations are captured once into a hipGraph and replayed until
// every system has converged.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "gis_kernels.hpp"

namespace gis {

namespace {

struct BatchView {
    const int* tile_sys;     // tile -> system
    const int* tile0;        // system -> first tile
    KState* st;              // system -> state
    const RedWs* red;        // system -> reduction workspace (tile-local indexing)
};

struct SysTile {
    int sys, local;
    __device__ SysTile(const BatchView& bv, int tile) {
        sys = bv.tile_sys[tile];
        local = tile - bv.tile0[sys];
    }
};

template <class P>
__global__ __launch_bounds__(kWG) void k_bsetup(Mat<double> m, int N, const double* __restrict__ b,
                                                const double* __restrict__ x,
                                                double* __restrict__ r, BatchView bv) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const SysTile s(bv, t.tile);
    KState* st = &bv.st[s.sys];
    P::run(m, N, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    double bv4[4], rv[4];
    load4(b, t.i0, N, bv4);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rv[k] = bv4[k] - sh_y[threadIdx.x * kLeaf + k];
        if (t.i0 + k < N) leaf[0] = fma(rv[k], rv[k], leaf[0]);
    }
    store4(r, t.i0, N, rv);
    finish_tile_slots<1>(bv.red[s.sys], st, s.local, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        st->rho0 = v[0];
        st->rho1 = 0.0;
        st->norm = norm;
        st->norm0 = norm;
        st->k = 0;
        st->status = kOK;
        if (st->hist_cap > 0) st->hist[0] = norm;
        st->done = !(norm > st->error);
    });
}

template <class P>
__global__ __launch_bounds__(kWG) void k_bcg1(Mat<double> m, int N, const double* __restrict__ r,
                                              double* __restrict__ pb0, double* __restrict__ pb1,
                                              double* __restrict__ x, double* __restrict__ q,
                                              BatchView bv) {
    TileRows t(N);
    const SysTile s(bv, t.tile);
    KState* st = &bv.st[s.sys];
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    const int k = st->k + 1;
    const bool first = (k == 1);
    const double beta = st->beta;
    const double alpha_prev = st->alpha;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    if (first) {
        P::run(m, N, t.row0, t.nrows, [&](int c) { return r[c]; }, sh, sh_y);
    } else {
        P::run(m, N, t.row0, t.nrows, [&](int c) { return fma(beta, p_old[c], r[c]); }, sh,
               sh_y);
    }
    double rv[4], pv[4], xv[4], pn[4], qv[4];
    load4(r, t.i0, N, rv);
    if (!first) {
        load4(p_old, t.i0, N, pv);
        load4(x, t.i0, N, xv);
    }
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        pn[kk] = first ? rv[kk] : fma(beta, pv[kk], rv[kk]);
        qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
        if (!first) xv[kk] = fma(alpha_prev, pv[kk], xv[kk]);
        if (t.i0 + kk < N) leaf[0] = fma(pn[kk], qv[kk], leaf[0]);
    }
    store4(p_new, t.i0, N, pn);
    store4(q, t.i0, N, qv);
    if (!first) store4(x, t.i0, N, xv);
    finish_tile_slots<1>(bv.red[s.sys], st, s.local, leaf, shr,
                   [&](const double (&v)[1]) { st->alpha = st->rho0 / v[0]; });
}

// Split direction form (the schedule's fp64_direction = 1): bdir streams
// p_k = r (k = 1) or fma(beta, p_{k-1}, r) and the deferred x += alpha_{k-1} p_{k-1};
// bcg1s is then a pure SpMV of p_k with (p_k, q).  The same fma per row and the
// same leaves as k_bcg1, so the same bits.
__global__ __launch_bounds__(kWG) void k_bdir(int N, const double* __restrict__ r,
                                              double* __restrict__ pb0, double* __restrict__ pb1,
                                              double* __restrict__ x, BatchView bv) {
    TileRows t(N);
    const SysTile s(bv, t.tile);
    const KState* st = &bv.st[s.sys];
    if (st->done) return;
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    double rv[4];
    load4(r, t.i0, N, rv);
    if (k == 1) {
        store4(p_new, t.i0, N, rv);
        return;
    }
    const double beta = st->beta;
    const double alpha_prev = st->alpha;
    double pv[4], xv[4];
    load4(p_old, t.i0, N, pv);
    load4(x, t.i0, N, xv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        xv[kk] = fma(alpha_prev, pv[kk], xv[kk]);
        pv[kk] = fma(beta, pv[kk], rv[kk]);
    }
    store4(p_new, t.i0, N, pv);
    store4(x, t.i0, N, xv);
}

template <class P>
__global__ __launch_bounds__(kWG) void k_bcg1s(Mat<double> m, int N,
                                               const double* __restrict__ pb0,
                                               const double* __restrict__ pb1,
                                               double* __restrict__ q, BatchView bv) {
    TileRows t(N);
    const SysTile s(bv, t.tile);
    KState* st = &bv.st[s.sys];
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    const double* p = ((st->k + 1) & 1) ? pb1 : pb0;
    P::run(m, N, t.row0, t.nrows, [&](int c) { return p[c]; }, sh, sh_y);
    double pv[4], qv[4];
    load4(p, t.i0, N, pv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        qv[kk] = sh_y[threadIdx.x * kLeaf + kk];
        if (t.i0 + kk < N) leaf[0] = fma(pv[kk], qv[kk], leaf[0]);
    }
    store4(q, t.i0, N, qv);
    finish_tile_slots<1>(bv.red[s.sys], st, s.local, leaf, shr,
                         [&](const double (&v)[1]) { st->alpha = st->rho0 / v[0]; });
}

__global__ __launch_bounds__(kWG) void k_bcg2(int N, double* __restrict__ r,
                                              const double* __restrict__ q, BatchView bv) {
    TileRows t(N);
    const SysTile s(bv, t.tile);
    KState* st = &bv.st[s.sys];
    if (st->done) return;
    __shared__ RedShared shr;
    const double alpha = st->alpha;
    double rv[4], qv[4];
    load4(r, t.i0, N, rv);
    load4(q, t.i0, N, qv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rv[k] = fma(-alpha, qv[k], rv[k]);
        if (t.i0 + k < N) leaf[0] = fma(rv[k], rv[k], leaf[0]);
    }
    store4(r, t.i0, N, rv);
    finish_tile_slots<1>(bv.red[s.sys], st, s.local, leaf, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = st->rho0 / st->rho1;
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (!(norm > st->error) || st->k >= st->max_steps) st->done = 1;
    });
}

__global__ __launch_bounds__(kWG) void k_bcg_flush(int N, double* __restrict__ x,
                                                   const double* __restrict__ pb0,
                                                   const double* __restrict__ pb1,
                                                   BatchView bv) {
    TileRows t(N);
    const SysTile s(bv, t.tile);
    const KState* st = &bv.st[s.sys];
    const int k = st->k;
    if (k < 1) return;
    const double alpha = st->alpha;
    const double* p = (k & 1) ? pb1 : pb0;
    double xv[4], pv[4];
    load4(x, t.i0, N, xv);
    load4(p, t.i0, N, pv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) xv[kk] = fma(alpha, pv[kk], xv[kk]);
    store4(x, t.i0, N, xv);
}

// rows of system s in the padded layout <-> the caller's packed layout
__global__ __launch_bounds__(kWG) void k_pack(int S, const int* __restrict__ pad_off,
                                              const int* __restrict__ pack_off,
                                              const double* __restrict__ src,
                                              double* __restrict__ dst, int to_padded) {
    const int s = blockIdx.y;
    const int n = pack_off[s + 1] - pack_off[s];
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        if (to_padded) dst[pad_off[s] + i] = src[pack_off[s] + i];
        else dst[pack_off[s] + i] = src[pad_off[s] + i];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int batch_set_pattern(gis_context* h, int S, const int* n_cells, const int* n_faces,
                      const int* la, const int* ua) {
    if (S < 1) return fail(h, kErrArg, "need at least one system");
    BatchInfo& B = h->batch;
    B.ready = false;
    B.graph_valid = false;
    if (B.graph_exec) (void)hipGraphExecDestroy(B.graph_exec);
    B.graph_exec = nullptr;
    h->n_cpl = 0;   // the stacked systems carry no interface entries
    B.S = S;
    B.n.assign(n_cells, n_cells + S);
    B.m.assign(n_faces, n_faces + S);
    B.pad_off.resize(S + 1);
    B.pack_off.resize(S + 1);
    B.face_off.resize(S + 1);
    B.tile0.resize(S + 1);
    long long rows = 0, packed = 0, faces = 0;
    for (int s = 0; s < S; ++s) {
        if (n_cells[s] < 1 || n_faces[s] < 0) return fail(h, kErrArg, "bad system size");
        B.pad_off[s] = (int)rows;
        B.pack_off[s] = (int)packed;
        B.face_off[s] = (int)faces;
        B.tile0[s] = (int)(rows / kTile);
        rows += (long long)tiles_of(n_cells[s]) * kTile;
        packed += n_cells[s];
        faces += n_faces[s];
        if (rows + 2 * faces >= (1LL << 31)) return fail(h, kErrShape, "batch too large");
    }
    B.pad_off[S] = (int)rows;
    B.pack_off[S] = (int)packed;
    B.face_off[S] = (int)faces;
    B.tile0[S] = (int)(rows / kTile);
    // global padded addressing
    std::vector<int> gla((size_t)faces), gua((size_t)faces);
    for (int s = 0; s < S; ++s)
        for (int f = 0; f < n_faces[s]; ++f) {
            const size_t g = (size_t)B.face_off[s] + f;
            const int l = la[g], u = ua[g];
            if (l < 0 || l >= n_cells[s] || u < 0 || u >= n_cells[s])
                return fail(h, kErrArg, "face addressing out of range for its system");
            gla[g] = B.pad_off[s] + l;
            gua[g] = B.pad_off[s] + u;
        }
    const size_t bf = sizeof(int) * (size_t)faces;
    GIS_HIP(h, h->lower_addr.ensure(bf));
    GIS_HIP(h, h->upper_addr.ensure(bf));
    if (faces) {
        GIS_HIP(h, hipMemcpyAsync(h->lower_addr.p, gla.data(), bf, hipMemcpyHostToDevice,
                                  h->stream));
        GIS_HIP(h, hipMemcpyAsync(h->upper_addr.p, gua.data(), bf, hipMemcpyHostToDevice,
                                  h->stream));
    }
    GIS_TRY(build_pattern(h, (int)rows, (int)faces, h->lower_addr.as<int>(),
                          h->upper_addr.as<int>()));
    pattern_schedule(h);   // SellD when every slice of the stack has an offset table
    // per-system reduction workspaces and states
    const int T = B.tile0[S];
    std::vector<int> tile_sys(T);
    for (int s = 0; s < S; ++s)
        for (int t = B.tile0[s]; t < B.tile0[s + 1]; ++t) tile_sys[t] = s;
    std::vector<RedWs> reds(S);
    size_t dbl = 0;
    for (int s = 0; s < S; ++s) {
        int count = B.tile0[s + 1] - B.tile0[s];
        int L = 0;
        for (; L < kMaxLevels; ++L) {
            reds[s].count[L] = count;
            if (L >= 2) dbl += (size_t)count * kMaxNV;
            if (count == 1) break;
            count = (count + kGroup - 1) / kGroup;
        }
        reds[s].nlevels = L + 1;
        for (int j = L + 1; j < kMaxLevels; ++j) reds[s].count[j] = 1;
    }
    // levels 2 and up (the top collector's global path) as plain partials
    GIS_HIP(h, B.red_mem.ensure(dbl * sizeof(double) + 256));
    double* pp = B.red_mem.as<double>();
    for (int s = 0; s < S; ++s)
        for (int j = 0; j < kMaxLevels; ++j) {
            reds[s].part[j] = nullptr;
            if (j >= 2 && j < reds[s].nlevels) {
                reds[s].part[j] = pp;
                pp += (size_t)reds[s].count[j] * kMaxNV;
            }
        }
    // published-partial slots per system (level 0, then level 1)
    size_t nslot = 0;
    for (int s = 0; s < S; ++s) nslot += ((size_t)reds[s].count[0] + reds[s].count[1]) * kMaxNV;
    GIS_HIP(h, B.slot_mem.ensure(sizeof(double) * nslot));
    double* sp = B.slot_mem.as<double>();
    for (int s = 0; s < S; ++s) {
        reds[s].slot = sp;
        reds[s].slot1 = sp + (size_t)reds[s].count[0] * kMaxNV;
        sp += ((size_t)reds[s].count[0] + reds[s].count[1]) * kMaxNV;
    }
    B.slot_n = nslot;
    B.slots_dirty = false;
    GIS_TRY(fill_slots(h, B.slot_mem.as<double>(), nslot));
    GIS_HIP(h, B.meta.ensure(sizeof(RedWs) * S + sizeof(int) * ((size_t)T + 3 * (S + 1)) + 64));
    char* mb = B.meta.as<char>();
    B.d_red = reinterpret_cast<RedWs*>(mb);
    B.d_tile_sys = reinterpret_cast<int*>(mb + sizeof(RedWs) * S);
    B.d_tile0 = B.d_tile_sys + T;
    B.d_pad_off = B.d_tile0 + (S + 1);
    B.d_pack_off = B.d_pad_off + (S + 1);
    GIS_HIP(h, hipMemcpyAsync(B.d_red, reds.data(), sizeof(RedWs) * S, hipMemcpyHostToDevice,
                              h->stream));
    GIS_HIP(h, hipMemcpyAsync(B.d_tile_sys, tile_sys.data(), sizeof(int) * T,
                              hipMemcpyHostToDevice, h->stream));
    GIS_HIP(h, hipMemcpyAsync(B.d_tile0, B.tile0.data(), sizeof(int) * (S + 1),
                              hipMemcpyHostToDevice, h->stream));
    GIS_HIP(h, hipMemcpyAsync(B.d_pad_off, B.pad_off.data(), sizeof(int) * (S + 1),
                              hipMemcpyHostToDevice, h->stream));
    GIS_HIP(h, hipMemcpyAsync(B.d_pack_off, B.pack_off.data(), sizeof(int) * (S + 1),
                              hipMemcpyHostToDevice, h->stream));
    GIS_HIP(h, B.st.ensure(sizeof(KState) * S));
    if (B.st_host) (void)hipHostFree(B.st_host);
    B.st_host = nullptr;
    void* p = nullptr;
    GIS_HIP(h, hipHostMalloc(&p, sizeof(KState) * S, hipHostMallocDefault));
    B.st_host = static_cast<KState*>(p);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    B.ready = true;
    B.graph_valid = false;
    return kOK;
}

int batch_set_values(gis_context* h, const double* diag, const double* lower,
                     const double* upper, bool dev) {
    BatchInfo& B = h->batch;
    if (!B.ready) return fail(h, kErrState, "batch pattern not set");
    const int S = B.S;
    const size_t rows = (size_t)B.pad_off[S], packed = (size_t)B.pack_off[S];
    const size_t faces = (size_t)B.face_off[S];
    GIS_HIP(h, h->diag.ensure(sizeof(double) * rows));
    GIS_HIP(h, h->lower.ensure(sizeof(double) * std::max<size_t>(faces, 1)));
    GIS_HIP(h, h->upper.ensure(sizeof(double) * std::max<size_t>(faces, 1)));
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (faces) {   // faces keep their packed order (system by system)
        GIS_HIP(h, hipMemcpyAsync(h->lower.p, lower, sizeof(double) * faces, kin, h->stream));
        GIS_HIP(h, hipMemcpyAsync(h->upper.p, upper, sizeof(double) * faces, kin, h->stream));
    }
    // diagonal: packed -> padded, padding rows stay 0
    GIS_HIP(h, hipMemsetAsync(h->diag.p, 0, sizeof(double) * rows, h->stream));
    GIS_HIP(h, h->io2.ensure(sizeof(double) * packed));
    GIS_HIP(h, hipMemcpyAsync(h->io2.p, diag, sizeof(double) * packed, kin, h->stream));
    hipLaunchKernelGGL(k_pack, dim3(64, S), dim3(kWG), 0, h->stream, S, B.d_pad_off,
                       B.d_pack_off, h->io2.as<double>(), h->diag.as<double>(), 1);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(gather_values(h));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return kOK;
}

// one CG iteration of every system: the fused direction form, or the split one
static void launch_iteration(hipStream_t s, bool selld, bool split, const Mat<double>& m, int N,
                             double* r, double* pb0, double* pb1, double* x, double* q,
                             const BatchView& bv) {
    const dim3 g = tiles_grid(N);
    if (split) {
        hipLaunchKernelGGL(k_bdir, g, dim3(kWG), 0, s, N, r, pb0, pb1, x, bv);
        if (selld)
            hipLaunchKernelGGL((k_bcg1s<SellD>), g, dim3(kWG), 0, s, m, N, pb0, pb1, q, bv);
        else
            hipLaunchKernelGGL((k_bcg1s<Sell>), g, dim3(kWG), 0, s, m, N, pb0, pb1, q, bv);
    } else if (selld) {
        hipLaunchKernelGGL((k_bcg1<SellD>), g, dim3(kWG), 0, s, m, N, r, pb0, pb1, x, q, bv);
    } else {
        hipLaunchKernelGGL((k_bcg1<Sell>), g, dim3(kWG), 0, s, m, N, r, pb0, pb1, x, q, bv);
    }
    hipLaunchKernelGGL(k_bcg2, g, dim3(kWG), 0, s, N, r, q, bv);
}

static int batch_graph(gis_context* h, int iters, const Mat<double>& m, int N, double* r,
                       double* pb0, double* pb1, double* x, double* q, const BatchView& bv) {
    BatchInfo& B = h->batch;
    const bool selld = h->sched.spmv64 == SpmvKind::SellD;
    const bool split = h->sched.fp64_direction != 0;
    if (B.graph_valid && B.graph_iters == iters && B.graph_key == (void*)x &&
        B.graph_selld == selld && B.graph_split == split)
        return kOK;
    if (B.graph_exec) (void)hipGraphExecDestroy(B.graph_exec);
    B.graph_exec = nullptr;
    if (!B.cap_stream) GIS_HIP(h, hipStreamCreateWithFlags(&B.cap_stream, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    GIS_HIP(h, hipStreamBeginCapture(B.cap_stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < iters; ++i) launch_iteration(B.cap_stream, selld, split, m, N, r, pb0,
                                                     pb1, x, q, bv);
    hipError_t e = hipStreamEndCapture(B.cap_stream, &graph);
    if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&B.graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
    B.graph_valid = true;
    B.graph_iters = iters;
    B.graph_key = (void*)x;
    B.graph_selld = selld;
    B.graph_split = split;
    return kOK;
}

int batch_solve(gis_context* h, double* x_packed, const double* b_packed, bool dev,
                const double* error, long long max_steps, int use_graph, int* iters,
                double* final_res, int* status) {
    BatchInfo& B = h->batch;
    if (!B.ready || !h->has_values) return fail(h, kErrState, "batch pattern/values not set");
    const int S = B.S;
    const int N = h->N;   // padded rows
    const size_t n = (size_t)N, packed = (size_t)B.pack_off[S];
    GIS_HIP(h, h->vec.ensure(sizeof(double) * n * 6 + sizeof(double) * packed * 2));
    double* x = h->vec.as<double>();
    double* b = x + n;
    double* r = b + n;
    double* pb0 = r + n;
    double* pb1 = pb0 + n;
    double* q = pb1 + n;
    double* stage = q + n;   // packed staging (2 * packed)
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(stage, x_packed, sizeof(double) * packed, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(stage + packed, b_packed, sizeof(double) * packed, kin, h->stream));
    GIS_HIP(h, hipMemsetAsync(x, 0, sizeof(double) * n * 2, h->stream));
    hipLaunchKernelGGL(k_pack, dim3(64, S), dim3(kWG), 0, h->stream, S, B.d_pad_off,
                       B.d_pack_off, stage, x, 1);
    hipLaunchKernelGGL(k_pack, dim3(64, S), dim3(kWG), 0, h->stream, S, B.d_pad_off,
                       B.d_pack_off, stage + packed, b, 1);
    GIS_HIP(h, hipGetLastError());
    // states
    const long long ms = max_steps <= 0 ? (long long)INT_MAX : max_steps;
    for (int s = 0; s < S; ++s) {
        KState st{};
        st.error = error[s];
        st.max_steps = ms;
        st.hist = nullptr;
        st.hist_cap = 0;
        st.slot_fault = h->slot_fault;
        B.st_host[s] = st;
    }
    GIS_HIP(h, hipMemcpyAsync(B.st.p, B.st_host, sizeof(KState) * S, hipMemcpyHostToDevice,
                              h->stream));
    if (B.slots_dirty) {
        GIS_TRY(fill_slots(h, B.slot_mem.as<double>(), B.slot_n));
        B.slots_dirty = false;
    }
    const BatchView bv{B.d_tile_sys, B.d_tile0, B.st.as<KState>(), B.d_red};
    const Mat<double> m = mat64(h);
    const dim3 g = tiles_grid(N);
    hipLaunchKernelGGL((k_bsetup<Sell>), g, dim3(kWG), 0, h->stream, m, N, b, x, r, bv);
    GIS_HIP(h, hipGetLastError());
    const int block = std::max(1, std::min(h->sched.block_iters, 32));
    if (use_graph) GIS_TRY(batch_graph(h, block, m, N, r, pb0, pb1, x, q, bv));
    hipEvent_t ev_main = nullptr, ev_cap = nullptr;
    if (use_graph) {
        GIS_HIP(h, hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
        GIS_HIP(h, hipEventCreateWithFlags(&ev_cap, hipEventDisableTiming));
    }
    long long launched = 0;
    int rc = kOK;
    for (;;) {
        if (use_graph) {
            // order the graph after the work already on the handle stream, and back
            if (hipEventRecord(ev_main, h->stream) != hipSuccess ||
                hipStreamWaitEvent(B.cap_stream, ev_main, 0) != hipSuccess ||
                hipGraphLaunch(B.graph_exec, B.cap_stream) != hipSuccess ||
                hipEventRecord(ev_cap, B.cap_stream) != hipSuccess ||
                hipStreamWaitEvent(h->stream, ev_cap, 0) != hipSuccess) {
                rc = fail(h, kErrHip, "graph replay failed");
                break;
            }
        } else {
            for (int i = 0; i < block; ++i)
                launch_iteration(h->stream, h->sched.spmv64 == SpmvKind::SellD,
                                 h->sched.fp64_direction != 0, m, N, r, pb0, pb1, x, q, bv);
        }
        launched += block;
        if (hipMemcpyAsync(B.st_host, B.st.p, sizeof(KState) * S, hipMemcpyDeviceToHost,
                           h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail(h, kErrHip, "batch state read failed");
            break;
        }
        bool all = true;
        for (int s = 0; s < S; ++s) all = all && B.st_host[s].done;
        if (all || launched >= ms) break;
    }
    if (ev_main) (void)hipEventDestroy(ev_main);
    if (ev_cap) (void)hipEventDestroy(ev_cap);
    if (rc != kOK) return rc;
    hipLaunchKernelGGL(k_bcg_flush, g, dim3(kWG), 0, h->stream, N, x, pb0, pb1, bv);
    hipLaunchKernelGGL(k_pack, dim3(64, S), dim3(kWG), 0, h->stream, S, B.d_pad_off,
                       B.d_pack_off, x, stage, 0);
    GIS_HIP(h, hipGetLastError());
    GIS_HIP(h, hipMemcpyAsync(x_packed, stage, sizeof(double) * packed, kout, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    int worst = kOK;
    for (int s = 0; s < S; ++s) {
        const KState& st = B.st_host[s];
        int sk = st.status;
        if (sk == kErrSync) B.slots_dirty = true;
        if (sk == kOK && !(st.norm == st.norm)) sk = kNaN;
        if (iters) iters[s] = st.k;
        if (final_res) final_res[s] = st.norm;
        if (status) status[s] = sk;
        worst = std::max(worst, sk);
    }
    return worst;
}

void batch_release(gis_context* h) {
    BatchInfo& B = h->batch;
    if (B.graph_exec) (void)hipGraphExecDestroy(B.graph_exec);
    B.graph_exec = nullptr;
    if (B.cap_stream) (void)hipStreamDestroy(B.cap_stream);
    B.cap_stream = nullptr;
    if (B.st_host) (void)hipHostFree(B.st_host);
    B.st_host = nullptr;
}

}  // namespace gis
