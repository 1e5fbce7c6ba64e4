// Krylov / stationary solvers of gis on gfx950, FP64.
//
// Each method keeps its scalars (rho, alpha, ...) in a device KState and
// advances with a fixed sequence of launches per iteration; reductions finish
// inside the producing kernel (last-arriving workgroup), so no iteration
// needs a host round-trip.  The host enqueues iterations in blocks and only
// reads the `done` flag between blocks; launches after convergence return at
// once, which preserves the reference's per-iteration stopping rule and
// therefore its iteration count.
//
// Methods and the reference loops they follow (gis/SparseSystem/SolverSingleGPU.h):
//   CG :366-445   BiCG :568-675   Richardson :762-819 (+OptimizePower :723-752)
//   MR :825-890   Jacobi :924-994  MRJacobi :1000-1073  JMR :1079-1155  MRS :1236-1303
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "gis_internal.hpp"

namespace gis {

// ---------------------------------------------------------------------------
// SpMV schedules
// ---------------------------------------------------------------------------
template <int CAP>
struct Staged {
    template <typename T>
    using Shared = SpmvShared<T, CAP>;
    template <typename T, class XOp>
    __device__ static void run(const int* rowptr, const int* col, const T* val, int row0,
                               int nrows, XOp&& xop, Shared<T>& sh, T* y) {
        tile_spmv<T, CAP>(rowptr, col, val, row0, nrows, xop, sh, y);
    }
};

struct Direct {
    template <typename T>
    struct Shared {};
    template <typename T, class XOp>
    __device__ static void run(const int* rowptr, const int* col, const T* val, int row0,
                               int nrows, XOp&& xop, Shared<T>&, T* y) {
        tile_spmv_direct<T>(rowptr, col, val, row0, nrows, xop, y);
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
        default: {                                    \
            using POLICY = Direct;                    \
            __VA_ARGS__;                              \
        } break;                                      \
    }

// row range of this workgroup's tile
struct TileRows {
    int tile, row0, nrows, i0;   // i0: first of this thread's kLeaf rows
    __device__ TileRows(int N) {
        tile = blockIdx.x;
        row0 = tile * kTile;
        nrows = min(kTile, N - row0);
        i0 = row0 + threadIdx.x * kLeaf;
    }
};

// ---------------------------------------------------------------------------
// plain SpMV  y = A x  (SpMvKernel, kernels.h:206-237)
// ---------------------------------------------------------------------------
template <class P, typename T>
__global__ __launch_bounds__(kWG) void k_spmv(const int* __restrict__ rowptr,
                                              const int* __restrict__ col,
                                              const T* __restrict__ val, int N,
                                              const T* __restrict__ x, T* __restrict__ y) {
    __shared__ typename P::template Shared<T> sh;
    __shared__ T sh_y[kTile];
    TileRows t(N);
    P::run(rowptr, col, val, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    T o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = sh_y[threadIdx.x * kLeaf + k];
    store4(y, t.i0, N, o);
}

// ---------------------------------------------------------------------------
// reductions into KState::aux[slot]
// ---------------------------------------------------------------------------
// aux[slot] = dot(a, b)    (DotKernel, kernels.h:153-199)
__global__ __launch_bounds__(kWG) void k_dot(int N, const double* __restrict__ a,
                                             const double* __restrict__ b, KState* st,
                                             RedWs red, int slot) {
    __shared__ RedShared shr;
    TileRows t(N);
    double av[4], bv[4];
    load4(a, t.i0, N, av);
    load4(b, t.i0, N, bv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = fma(av[k], bv[k], leaf[0]);
    finish_tile<1>(red, t.tile, leaf, shr, [&](const double (&v)[1]) { st->aux[slot] = v[0]; });
}

// aux[0] = sum(x) / N   (ReductionKernel kernels.h:100-145 + gis.cu:108)
__global__ __launch_bounds__(kWG) void k_mean(int N, const double* __restrict__ x, KState* st,
                                              RedWs red) {
    __shared__ RedShared shr;
    TileRows t(N);
    double xv[4];
    load4(x, t.i0, N, xv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = leaf[0] + xv[k];
    finish_tile<1>(red, t.tile, leaf, shr, [&](const double (&v)[1]) { st->aux[0] = v[0] / N; });
}

// NormFactor (gis/gis.cu:106-123):
//   aux[1] = ||A x - A x_bar||_2 + ||b - A x_bar||_2 + 1e-20, with x_bar = aux[0]
template <class P>
__global__ __launch_bounds__(kWG) void k_norm_factor(const int* __restrict__ rowptr,
                                                     const int* __restrict__ col,
                                                     const double* __restrict__ val, int N,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ x, KState* st,
                                                     RedWs red) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ double sh_yb[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const double x_bar = st->aux[0];
    P::run(rowptr, col, val, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
    P::run(rowptr, col, val, t.row0, t.nrows, [&](int) { return x_bar; }, sh, sh_yb);
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
    finish_tile<2>(red, t.tile, leaf, shr, [&](const double (&v)[2]) {
        st->aux[1] = sqrt(v[0]) + sqrt(v[1]) + 1e-20;
    });
}

// ---------------------------------------------------------------------------
// solver set-up: r = b - A x (ResidualKernel, kernels.h:409-441), optionally
// r = s o r (JMR, SolverSingleGPU.h:1098-1100), copies of r, and
// rho0 = (r, r), norm0 = sqrt(rho0), hist[0].
// done_rule 0: done when !(norm > error)   (CG, Jacobi, MR, ... loop guards)
// done_rule 1: never done before the first iteration (BiCG, :621)
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_setup(const int* __restrict__ rowptr,
                                               const int* __restrict__ col,
                                               const double* __restrict__ val, int N,
                                               const double* __restrict__ b,
                                               const double* __restrict__ x,
                                               const double* __restrict__ scale,
                                               double* __restrict__ r, double* __restrict__ r2,
                                               KState* st, RedWs red, int done_rule) {
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    P::run(rowptr, col, val, t.row0, t.nrows, [&](int c) { return x[c]; }, sh, sh_y);
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
    finish_tile<1>(red, t.tile, leaf, shr, [&](const double (&v)[1]) {
        const double norm = sqrt(v[0]);
        st->rho0 = v[0];
        st->rho1 = 0.0;
        st->norm = norm;
        st->norm0 = norm;
        st->k = 0;
        st->status = kOK;
        if (st->hist_cap > 0) st->hist[0] = norm;
        st->done = (done_rule == 0) ? !(norm > st->error) : 0;
    });
}

__device__ __forceinline__ void record_norm(KState* st, double norm) {
    const int k = st->k + 1;
    st->k = k;
    st->norm = norm;
    if (k < st->hist_cap) st->hist[k] = norm;
}

// ---------------------------------------------------------------------------
// CG  (SolverSingleGPU.h:366-445), two launches per iteration:
//   cg1: p_k = r + beta p_{k-1} (k > 1; formed on the fly, also at the gathered
//        columns), q = A p_k, x += alpha_{k-1} p_{k-1} (deferred), (p_k, q)
//        -> alpha_k = rho / (p_k, q)
//   cg2: r -= alpha_k q, rho = (r, r), norm, history, beta, stopping rule
// p alternates between two buffers (p_k in buffer k&1).
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kWG) void k_cg1(const int* __restrict__ rowptr,
                                             const int* __restrict__ col,
                                             const double* __restrict__ val, int N,
                                             const double* __restrict__ r,
                                             double* __restrict__ pb0, double* __restrict__ pb1,
                                             double* __restrict__ x, double* __restrict__ q,
                                             KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<double> sh;
    __shared__ double sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    const int k = st->k + 1;
    const bool first = (k == 1);
    const double beta = st->beta;
    const double alpha_prev = st->alpha;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    if (first) {
        P::run(rowptr, col, val, t.row0, t.nrows, [&](int c) { return r[c]; }, sh, sh_y);
    } else {
        P::run(rowptr, col, val, t.row0, t.nrows,
               [&](int c) { return fma(beta, p_old[c], r[c]); }, sh, sh_y);
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
    finish_tile<1>(red, t.tile, leaf, shr,
                   [&](const double (&v)[1]) { st->alpha = st->rho0 / v[0]; });
}

__global__ __launch_bounds__(kWG) void k_cg2(int N, double* __restrict__ r,
                                             const double* __restrict__ q, KState* st,
                                             RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
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
    finish_tile<1>(red, t.tile, leaf, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = st->rho0 / st->rho1;
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (!(norm > st->error) || st->k >= st->max_steps) st->done = 1;
    });
}

// x += alpha_k p_k for the last iteration (the update cg1 defers)
__global__ __launch_bounds__(kWG) void k_cg_flush(int N, double* __restrict__ x,
                                                  const double* __restrict__ pb0,
                                                  const double* __restrict__ pb1,
                                                  const KState* st) {
    const int k = st->k;
    if (k < 1) return;
    const double alpha = st->alpha;
    const double* p = (k & 1) ? pb1 : pb0;
    TileRows t(N);
    double xv[4], pv[4];
    load4(x, t.i0, N, xv);
    load4(p, t.i0, N, pv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) xv[kk] = fma(alpha, pv[kk], xv[kk]);
    store4(x, t.i0, N, xv);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// The level counts must describe exactly `rows` (the group-final logic waits
// for count[L] arrivals); memory only grows.  Tickets are re-zeroed whenever
// the shape changes.
int ensure_reduction(gis_context* h, int rows) {
    if (h->red_rows == rows && h->red_mem.p) return kOK;
    RedWs r{};
    int count = tiles_of(rows);
    size_t dbl = 0, ints = 0;
    int L = 0;
    for (; L < kMaxLevels; ++L) {
        r.count[L] = count;
        dbl += (size_t)count * kMaxNV;
        ints += (size_t)(count + kGroup - 1) / kGroup;
        if (count == 1) break;
        count = (count + kGroup - 1) / kGroup;
    }
    if (L == kMaxLevels) return fail(h, kErrShape, "reduction depth exceeded");
    r.nlevels = L + 1;
    for (int j = L + 1; j < kMaxLevels; ++j) r.count[j] = 1;
    GIS_HIP(h, h->red_mem.ensure(dbl * sizeof(double) + ints * sizeof(unsigned) + 256));
    char* base = h->red_mem.as<char>();
    size_t off = 0;
    for (int j = 0; j <= L; ++j) {
        r.part[j] = reinterpret_cast<double*>(base + off);
        off += (size_t)r.count[j] * kMaxNV * sizeof(double);
    }
    unsigned* tk = reinterpret_cast<unsigned*>(base + off);
    GIS_HIP(h, hipMemsetAsync(tk, 0, ints * sizeof(unsigned), h->stream));
    for (int j = 0; j <= L; ++j) {
        r.ticket[j] = tk;
        tk += (r.count[j] + kGroup - 1) / kGroup;
    }
    for (int j = L + 1; j < kMaxLevels; ++j) {
        r.part[j] = nullptr;
        r.ticket[j] = nullptr;
    }
    h->red = r;
    h->red_rows = rows;
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
    return kOK;
}

static int push_state(gis_context* h) {
    GIS_HIP(h, hipMemcpyAsync(h->st, h->st_host, sizeof(KState), hipMemcpyHostToDevice,
                              h->stream));
    return kOK;
}

static inline dim3 tiles_grid(int N) { return dim3((unsigned)tiles_of(N)); }

int launch_spmv(gis_context* h, const double* x, double* y) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_spmv<P, double>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, h->rowptr.as<int>(), h->col.as<int>(),
                                         h->val.as<double>(), h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int launch_spmv_t(gis_context* h, const double* x, double* y) {
    GIS_TRY(ensure_valT(h));
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_spmv<P, double>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, h->rowptr.as<int>(), h->col.as<int>(),
                                         h->valT.as<double>(), h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int launch_spmv32(gis_context* h, const float* x, float* y) {
    GIS_TRY(ensure_val32(h));
    GIS_SPMV_DISPATCH(h->sched.spmv32, P,
                      hipLaunchKernelGGL((k_spmv<P, float>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, h->rowptr.as<int>(), h->col.as<int>(),
                                         h->val32.as<float>(), h->N, x, y));
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int device_dot(gis_context* h, const double* a, const double* b, double* out) {
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    hipLaunchKernelGGL(k_dot, tiles_grid(h->N), dim3(kWG), 0, h->stream, h->N, a, b, h->st,
                       h->red, 7);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(read_state(h));
    *out = h->st_host->aux[7];
    return kOK;
}

int device_norm_factor(gis_context* h, const double* d_b, const double* d_x, double* nf) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    GIS_TRY(ensure_reduction(h, h->N));
    GIS_TRY(ensure_state(h));
    hipLaunchKernelGGL(k_mean, tiles_grid(h->N), dim3(kWG), 0, h->stream, h->N, d_x, h->st,
                       h->red);
    GIS_HIP(h, hipGetLastError());
    GIS_SPMV_DISPATCH(h->sched.spmv64, P,
                      hipLaunchKernelGGL((k_norm_factor<P>), tiles_grid(h->N), dim3(kWG), 0,
                                         h->stream, h->rowptr.as<int>(), h->col.as<int>(),
                                         h->val.as<double>(), h->N, d_b, d_x, h->st, h->red));
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(read_state(h));
    *nf = h->st_host->aux[1];
    return kOK;
}

// Enqueue iterations in blocks until the device reports done.
template <class Iter>
static int run_blocks(gis_context* h, long long max_steps, Iter&& launch_iteration) {
    long long launched = 0;
    int block = 4;
    for (;;) {
        for (int i = 0; i < block && launched < max_steps; ++i, ++launched)
            GIS_TRY(launch_iteration());
        GIS_HIP(h, hipGetLastError());
        GIS_TRY(read_state(h));
        if (h->st_host->done || launched >= max_steps) break;
        block = std::min(block * 2, std::max(4, h->sched.block_iters));
    }
    return kOK;
}

static int prepare(gis_context* h, const SolveArgs& a, int nvec) {
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
    s.hist = h->hist.as<double>();
    s.hist_cap = cap;
    *h->st_host = s;
    return push_state(h);
}

static int solve_cg(gis_context* h, const SolveArgs& a, double* x, const double* b) {
    GIS_TRY(prepare(h, a, 4));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* pb0 = r + N;
    double* pb1 = r + 2 * (size_t)N;
    double* q = r + 3 * (size_t)N;
    const dim3 g = tiles_grid(N);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        hipLaunchKernelGGL((k_setup<P>), g, dim3(kWG), 0, h->stream, h->rowptr.as<int>(),
                           h->col.as<int>(), h->val.as<double>(), N, b, x, nullptr, r, nullptr,
                           h->st, h->red, 0);
        GIS_HIP(h, hipGetLastError());
        const long long ms = a.max_steps <= 0 ? (long long)INT_MAX : a.max_steps;
        GIS_TRY(run_blocks(h, ms, [&]() -> int {
            hipLaunchKernelGGL((k_cg1<P>), g, dim3(kWG), 0, h->stream, h->rowptr.as<int>(),
                               h->col.as<int>(), h->val.as<double>(), N, r, pb0, pb1, x, q,
                               h->st, h->red);
            hipLaunchKernelGGL(k_cg2, g, dim3(kWG), 0, h->stream, N, r, q, h->st, h->red);
            return kOK;
        }));
    });
    hipLaunchKernelGGL(k_cg_flush, g, dim3(kWG), 0, h->stream, N, x, pb0, pb1, h->st);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int device_solve(gis_context* h, const SolveArgs& a, double* d_x, const double* d_b,
                 int* hist_len, double* norm0, double* norm_final) {
    int st;
    switch (a.method) {
        case 0: st = solve_cg(h, a, d_x, d_b); break;
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
