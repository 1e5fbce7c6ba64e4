// Kernel plumbing shared by the solver translation units: the matrix view,
// the SpMV schedules (all bit-identical), tile bookkeeping and the solver
// set-up kernel.
#pragma once

#include <type_traits>

#include "gis_internal.hpp"

namespace gis {

// Device view of one matrix in both layouts (CSR + sliced); a schedule reads
// whichever it needs.
template <typename T>
struct Mat {
    const int* rowptr;
    const int* col;
    const T* val;
    const int* soff;
    const int* scol;
    const T* sval;
    const signed char* sres;      // compressed slot columns (SellC)
    const int* sbase;
    const unsigned char* sflag;
    const int* ooff;              // column-offset tables (SellD)
    const int* doffs;
    const unsigned* dmask;
    const T* dval;
    const unsigned char* dmask8;  // row masks of slices with at most 8 table entries
    // symmetric storage (SellD of a bitwise symmetric matrix, every slice tabled):
    // per table entry {offset, value bases}, and the kept slot values; null: full slots
    const int4* symx;
    const T* dsv;
};

// ---------------------------------------------------------------------------
// SpMV schedules: y[row0 .. row0+nrows) -> sh_y (tile row order)
// ---------------------------------------------------------------------------
struct Sell {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int N, int row0, int, XOp&& xop, Shared<T>&,
                               T* y) {
        tile_sell<T>(m.soff, m.scol, m.sval, N, row0, xop, y);
    }
};

// Sliced rows with compressed columns (1 byte per slot instead of 4).
struct SellC {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int N, int row0, int, XOp&& xop, Shared<T>&,
                               T* y) {
        tile_sell_c<T>(m.soff, m.scol, m.sres, m.sbase, m.sflag, m.sval, N, row0, xop, y);
    }
};

// Sliced rows with per-slice column-offset tables and row masks, two slices at
// a time (no column index per slot).
struct SellD {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int N, int row0, int, XOp&& xop, Shared<T>&,
                               T* y) {
        // the symmetric storage is FP64 only: FP32 kernels carry no second path
        if constexpr (std::is_same<T, double>::value) {
            if (m.dsv) {
                tile_sell_ds<T>(m.ooff, m.dmask, m.dmask8, m.symx, m.dsv, N, row0, xop, y);
                return;
            }
        }
        tile_sell_d<T>(m.soff, m.scol, m.sval, m.ooff, m.doffs, m.dmask, m.dmask8, m.dval, N,
                       row0, xop, y);
    }
};

// Sliced rows, two slices per lane at a time (more loads in flight per wave).
struct Sell2 {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int N, int row0, int, XOp&& xop, Shared<T>&,
                               T* y) {
        tile_sell2<T>(m.soff, m.scol, m.sval, N, row0, xop, y);
    }
};

template <int CAP>
struct Staged {
    template <typename T>
    using Shared = SpmvShared<T, CAP>;
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int, int row0, int nrows, XOp&& xop,
                               Shared<T>& sh, T* y) {
        tile_spmv<T, CAP>(m.rowptr, m.col, m.val, row0, nrows, xop, sh, y);
    }
};

struct Direct {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const Mat<T>& m, int, int row0, int nrows, XOp&& xop,
                               Shared<T>&, T* y) {
        tile_spmv_direct<T>(m.rowptr, m.col, m.val, row0, nrows, xop, y);
    }
};

#define GIS_SPMV_DISPATCH(kind, POLICY, ...)          \
    switch (kind) {                                   \
        case SpmvKind::Staged2K: {                    \
            using POLICY = Staged<2048>;              \
            __VA_ARGS__;                              \
        } break;                                      \
        case SpmvKind::Staged4K: {                    \
            using POLICY = Staged<4096>;              \
            __VA_ARGS__;                              \
        } break;                                      \
        case SpmvKind::Direct: {                      \
            using POLICY = Direct;                    \
            __VA_ARGS__;                              \
        } break;                                      \
        case SpmvKind::Sell2: {                       \
            using POLICY = Sell2;                     \
            __VA_ARGS__;                              \
        } break;                                      \
        case SpmvKind::SellC: {                       \
            using POLICY = SellC;                     \
            __VA_ARGS__;                              \
        } break;                                      \
        case SpmvKind::SellD: {                       \
            using POLICY = SellD;                     \
            __VA_ARGS__;                              \
        } break;                                      \
        default: {                                    \
            using POLICY = Sell;                      \
            __VA_ARGS__;                              \
        } break;                                      \
    }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (each
// with its own L2).  Map workgroup i to a tile so that XCD x sweeps one
// contiguous range of tiles: neighbouring rows -- and the x gathers of a
// tile -- then stay in one L2.  The map is a bijection on [0, G); results do
// not depend on it (partials are indexed by tile).
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_tile(int i, int G) {
    const int q = G / kXcds;
    const int r = G % kXcds;
    const int x = i % kXcds;
    const int j = i / kXcds;
    return x * q + min(x, r) + j;
}

// Row range of a workgroup tile.  TileRows(N): tile = workgroup index (best for
// the SpMV-carrying kernels, measured); StreamRows(N): XCD-contiguous tiles
// (best for the pure streaming update kernels, measured).
struct TileRows {
    int tile, row0, nrows, i0;   // i0: first of this thread's kLeaf rows
    __device__ TileRows(int N) : TileRows(N, (int)blockIdx.x) {}
    __device__ TileRows(int N, int tile_) {
        tile = tile_;
        row0 = tile * kTile;
        nrows = min(kTile, N - row0);
        i0 = row0 + threadIdx.x * kLeaf;
    }
};

struct StreamRows : TileRows {
    __device__ StreamRows(int N) : TileRows(N, xcd_tile((int)blockIdx.x, (int)gridDim.x)) {}
};

// Coarsened workgroups: workgroup w covers the C consecutive tiles w*C ..
// w*C+C-1 (stream = true: w XCD-swizzled first).  nt <= C tiles exist.
template <int C>
struct SuperTile {
    int tile0, nt;
    __device__ SuperTile(int N, bool stream) {
        const int s = stream ? xcd_tile((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
        tile0 = s * C;
        nt = min(C, tiles_of(N) - tile0);
    }
};

// History entry k: the residual norm, or (detail = 1, DetailType::Time of
// support.h:7-9) the seconds since the loop started, read from the device's
// constant-rate wall clock by the workgroup that finishes iteration k.
__device__ __forceinline__ void record_norm(KState* st, double norm) {
    const int k = st->k + 1;
    st->k = k;
    st->norm = norm;
    if (k < st->hist_cap)
        st->hist[k] = st->detail ? (double)(wall_clock64() - st->t0) * st->tick : norm;
}

// ---------------------------------------------------------------------------
// solver set-up: r = b - A x (ResidualKernel, kernels.h:409-441), optionally
// r = s o r (JMR, SolverSingleGPU.h:1098-1100), a copy of r, and
// rho0 = (r, r), norm0 = sqrt(rho0), hist[0].
// done_rule 0: done when !(norm > error)   (loop guards of CG, Jacobi, MR, ...)
// done_rule 1: never done before the first iteration (BiCG, :621)
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_setup(Mat<double> m, int N,
                                               const double* __restrict__ b,
                                               const double* __restrict__ x,
                                               const double* __restrict__ scale,
                                               double* __restrict__ r, double* __restrict__ r2,
                                               KState* st, RedWs red, int done_rule) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    double bv[4], rv[4], sv[4];
    load4(b, t.i0, N, bv);
    if (scale) load4(scale, t.i0, N, sv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rv[k] = bv[k] - sh_y[threadIdx.x * kLeaf + k];
        if (scale) rv[k] = sv[k] * rv[k];
        if (t.i0 + k < N) leaf[0] = fma(rv[k], rv[k], leaf[0]);
    }
    store4(r, t.i0, N, rv);
    if (r2) store4(r2, t.i0, N, rv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        st->rho0 = v[0];
        st->rho1 = 0.0;
        st->norm = norm;
        st->norm0 = norm;
        st->k = 0;
        st->status = kOK;
        if (st->hist_cap > 0) st->hist[0] = st->detail ? 0.0 : norm;
        st->t0 = wall_clock64();
        if (st->rel_tol > 0.0) st->error = fmax(st->error, st->rel_tol * norm);
        st->done = (done_rule == 0) ? (!(norm > st->error) && st->min_steps <= 0) : 0;
    });
}

// ---------------------------------------------------------------------------
// CG pieces shared by CG (gis_solvers.hip) and PCG (gis_amg.hip)
// ---------------------------------------------------------------------------
// Split schedule, pass 2: q = A p_k with one gathered vector, (p_k, q) -> alpha_k.
// KState::aux slot of alpha_{k-1} (written by k_cg1s, read by the paired x
// updates of k_cg_dir and k_cg_flush)
constexpr int kAlphaPrevSlot = 5;

template <class P, int C>
__global__ __launch_bounds__(kWG) void k_cg1s(Mat<double> m, int N,
                                              const double* __restrict__ pb0,
                                              const double* __restrict__ pb1,
                                              double* __restrict__ q, KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, false);
    const double* p = ((st->k + 1) & 1) ? pb1 : pb0;
    double leaf[C][1];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        leaf[c][0] = 0.0;
        if (c >= sup.nt) continue;
        if (c > 0) __syncthreads();
        const TileRows t(N, sup.tile0 + c);
        P::run(m, N, t.row0, t.nrows, [&](int j) { return p[j]; }, sh, sh_y);
        double pv[4];
        load4(p, t.i0, N, pv);
        __syncthreads();
        store_tile_cached(q, sh_y, t.row0, N);   // k_cg2 reads q next
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            if (t.i0 + kk < N) leaf[c][0] = fma(pv[kk], sh_y[threadIdx.x * kLeaf + kk], leaf[c][0]);
    }
    SlotReduce<1, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
    tr.end(red, st, shr, [&](const double (&v)[1]) {
        st->aux[kAlphaPrevSlot] = st->alpha;   // alpha_{k-1}, for the paired x updates
        st->alpha = st->rho0 / v[0];
    });
}

namespace {

// x += alpha_k p_k for the last iteration (the update the direction pass
// defers).  paired (k_cg_dir): after an even iteration count K the update of
// iteration K - 1 is pending too, applied first: x = fma(alpha_K, p_K,
// fma(alpha_{K-1}, p_{K-1}, x)), the order of the per-iteration updates.
__global__ __launch_bounds__(kWG) void k_cg_flush(int N, double* __restrict__ x,
                                                  const double* __restrict__ pb0,
                                                  const double* __restrict__ pb1,
                                                  const KState* st, int paired) {
    const int k = st->k;
    if (k < 1) return;
    const double alpha = st->alpha;
    const double* p = (k & 1) ? pb1 : pb0;
    const bool two = paired && (k & 1) == 0;
    const double alpha_prev = st->aux[kAlphaPrevSlot];
    const double* pp = (k & 1) ? pb0 : pb1;
    StreamRows t(N);
    double xv[4], pv[4], ppv[4];
    load4(x, t.i0, N, xv);
    load4(p, t.i0, N, pv);
    if (two) load4(pp, t.i0, N, ppv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        if (two) xv[kk] = fma(alpha_prev, ppv[kk], xv[kk]);
        xv[kk] = fma(alpha, pv[kk], xv[kk]);
    }
    store4(x, t.i0, N, xv);
}

}  // namespace

// host helpers shared by the solver files
inline dim3 tiles_grid(int N) { return dim3((unsigned)tiles_of(N)); }
inline dim3 super_grid(int N, int C) { return dim3((unsigned)((tiles_of(N) + C - 1) / C)); }

// Tiles per workgroup of the iteration kernels (a schedule field: 1, 2 or 4).
#define GIS_COARSEN_DISPATCH(c, CC, ...)  \
    switch (c) {                          \
        case 4: {                         \
            constexpr int CC = 4;         \
            __VA_ARGS__;                  \
        } break;                          \
        case 2: {                         \
            constexpr int CC = 2;         \
            __VA_ARGS__;                  \
        } break;                          \
        default: {                        \
            constexpr int CC = 1;         \
            __VA_ARGS__;                  \
        } break;                          \
    }

inline Mat<double> mat64(const gis_context* h) {
    const bool sym = use_sym(h);
    return Mat<double>{h->rowptr.as<int>(), h->col.as<int>(), h->val.as<double>(),
                       h->soff.as<int>(), h->scol.as<int>(), h->sval.as<double>(),
                       h->sres.as<signed char>(), h->sbase.as<int>(),
                       h->sflag.as<unsigned char>(),
                       h->ooff.as<int>(), h->doffs.as<int>(), h->dmask.as<unsigned>(),
                       h->dval.as<double>(), h->dmask8.as<unsigned char>(),
                       sym ? h->symx.as<int4>() : nullptr, sym ? h->dsv.as<double>() : nullptr};
}

inline Mat<double> mat64T(const gis_context* h) {
    return Mat<double>{h->rowptr.as<int>(), h->col.as<int>(), h->valT.as<double>(),
                       h->soff.as<int>(), h->scol.as<int>(), h->svalT.as<double>(),
                       h->sres.as<signed char>(), h->sbase.as<int>(),
                       h->sflag.as<unsigned char>(),
                       h->ooff.as<int>(), h->doffs.as<int>(), h->dmask.as<unsigned>(),
                       h->dvalT.as<double>(), h->dmask8.as<unsigned char>()};
}

inline Mat<float> mat32T(const gis_context* h) {
    return Mat<float>{h->rowptr.as<int>(), h->col.as<int>(), h->val32T.as<float>(),
                      h->soff.as<int>(), h->scol.as<int>(), h->sval32T.as<float>(),
                      h->sres.as<signed char>(), h->sbase.as<int>(),
                      h->sflag.as<unsigned char>(),
                      h->ooff.as<int>(), h->doffs.as<int>(), h->dmask.as<unsigned>(),
                      h->dval32T.as<float>(), h->dmask8.as<unsigned char>()};
}

// FP32 keeps the full slots: its symmetric storage measured slower at 256^3
// (SpMV 103-116 us against 98 us, mixed inner iteration 228 against 203 us)
inline Mat<float> mat32(const gis_context* h) {
    return Mat<float>{h->rowptr.as<int>(), h->col.as<int>(), h->val32.as<float>(),
                      h->soff.as<int>(), h->scol.as<int>(), h->sval32.as<float>(),
                      h->sres.as<signed char>(), h->sbase.as<int>(),
                      h->sflag.as<unsigned char>(),
                      h->ooff.as<int>(), h->doffs.as<int>(), h->dmask.as<unsigned>(),
                      h->dval32.as<float>(), h->dmask8.as<unsigned char>(), nullptr, nullptr};
}

// Iterations the host may launch: the cap, or min_steps when that is larger
// (the kernels end the solve themselves through KState::done).
inline long long iteration_budget(const gis_context* h) {
    return h->st_host->max_steps > h->st_host->min_steps ? h->st_host->max_steps
                                                         : h->st_host->min_steps;
}

// Enqueue iterations in blocks until the device reports done.  predict:
// iterations are costly, so a block is cut to the iterations that the mean
// contraction so far says are left (plus 10%): fewer launches run after
// convergence.  Block sizes never change results (launches after `done` are
// no-ops).
template <class Iter>
int run_blocks(gis_context* h, long long max_steps, Iter&& launch_iteration,
               bool predict = false) {
    long long launched = 0;
    int block = 4;
    for (;;) {
        for (int i = 0; i < block && launched < max_steps; ++i, ++launched)
            GIS_TRY(launch_iteration());
        GIS_HIP(h, hipGetLastError());
        GIS_TRY(read_state(h));
        if (h->st_host->done || launched >= max_steps) break;
        block = block * 2 < h->sched.block_iters ? block * 2 : h->sched.block_iters;
        const KState& s = *h->st_host;
        if (predict && s.k >= 2 && s.norm > s.error && s.error > 0.0 && s.norm0 > s.norm &&
            s.k >= s.min_steps) {
            const double rate = std::log(s.norm / s.norm0) / s.k;
            const double left = std::log(s.error / s.norm) / rate;
            if (left > 0.0 && left * 1.1 + 1.0 < block) block = (int)(left * 1.1) + 1;
        }
        if (block < 1) block = 1;
    }
    return kOK;
}

int prepare_solve(gis_context* h, const SolveArgs& a, int nvec);

}  // namespace gis
