// Preconditioned CG on gfx950: preconditioner none, diagonal, or GAMG
// (pairwise algebraic agglomeration, Galerkin coarse LDU levels, symmetric
// V-cycle with damped Jacobi).  The reference has no preconditioners (its
// roadmap, README.md:71-73); the loop keeps the stopping rule and history of
// its CG (SolverSingleGPU.h:366-445), and with preconditioner none the
// iterates are those of gis CG, bit for bit.
//
// Per iteration (device state as in CG, no host round-trip):
//   dir:   p = z (k = 1) or fma(beta, p, z); deferred x = fma(alpha, p_prev, x)
//   spmv:  q = A p, alpha = rho / (p, q)                         (k_cg1s)
//   upd:   r = fma(-alpha, q, r), ||r||, history, stop test; diagonal: z = r / a_ii
//          and rho = (r, z) in the same pass
//   GAMG:  z = V-cycle(r); its last sweep also forms rho = (r, z)
//
// The hierarchy is built on the host once per pattern, from the values
// present at that time (agglomeration weights |upper| + |lower|): two
// pairwise passes per level until a level has <= kDense cells, whose dense
// inverse (Gauss-Jordan without pivoting) is formed on the device.  The
// Galerkin values of every coarse level are recomputed on the device after
// each change of values, in a fixed order (members ascending, then faces
// ascending), so results do not depend on the launch.
#include <algorithm>
#include <climits>
#include <cmath>
#include <exception>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "gis_agglom.hpp"
#include "gis_kernels.hpp"

namespace gis {

int solve_pcg(gis_context* h, const SolveArgs& a, int pc, double* x, const double* b);

namespace {

constexpr int kBlk = 256;
constexpr int kDense = 64;        // the last level is inverted densely at or below this size
constexpr int kAmgLevels = 32;
constexpr int kGammaSlot = 7;     // KState::aux slot of (r, z_prev) for GAMG32's beta

inline dim3 grid_of(long long n) { return dim3((unsigned)((n + kBlk - 1) / kBlk)); }

// ---------------------------------------------------------------------------
// device: coarse values
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlk) void k_amg_cdiag(int nc, const int* __restrict__ mptr,
                                                    const int* __restrict__ mlist,
                                                    const int* __restrict__ iptr,
                                                    const int* __restrict__ ilist,
                                                    const double* __restrict__ diag,
                                                    const double* __restrict__ lower,
                                                    const double* __restrict__ upper,
                                                    double* __restrict__ cdiag) {
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= nc) return;
    double s = 0.0;
    for (int p = mptr[c]; p < mptr[c + 1]; ++p) s = s + diag[mlist[p]];
    for (int p = iptr[c]; p < iptr[c + 1]; ++p) {
        const int f = ilist[p];
        s = s + upper[f];
        s = s + lower[f];
    }
    cdiag[c] = s;
}

__global__ __launch_bounds__(kBlk) void k_amg_cface(int mc, const int* __restrict__ fptr,
                                                    const int* __restrict__ flist,
                                                    const double* __restrict__ lower,
                                                    const double* __restrict__ upper,
                                                    double* __restrict__ clower,
                                                    double* __restrict__ cupper) {
    const int g = blockIdx.x * kBlk + threadIdx.x;
    if (g >= mc) return;
    double su = 0.0, sl = 0.0;
    for (int p = fptr[g]; p < fptr[g + 1]; ++p) {
        const int code = flist[p];
        const int f = code >> 1;
        const bool flip = code & 1;
        su = su + (flip ? lower[f] : upper[f]);
        sl = sl + (flip ? upper[f] : lower[f]);
    }
    cupper[g] = su;
    clower[g] = sl;
}

__global__ __launch_bounds__(kBlk) void k_amg_sm(int n, double omega,
                                                 const double* __restrict__ diag,
                                                 double* __restrict__ sm) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) sm[i] = omega / diag[i];
}

__global__ __launch_bounds__(kBlk) void k_recip(int n, const double* __restrict__ diag,
                                                double* __restrict__ out) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) out[i] = 1.0 / diag[i];
}

// Dense inverse of the last level by Gauss-Jordan without pivoting on the
// augmented [A | I] (n x 2n in M); thread j owns column j.
__global__ __launch_bounds__(2 * kDense) void k_amg_dense_inv(int n, int m,
                                                              const int* __restrict__ la,
                                                              const int* __restrict__ ua,
                                                              const double* __restrict__ diag,
                                                              const double* __restrict__ lower,
                                                              const double* __restrict__ upper,
                                                              double* __restrict__ M,
                                                              double* __restrict__ ainv) {
    __shared__ double fac[kDense];
    const int w = 2 * n;
    const int j = threadIdx.x;
    for (int idx = j; idx < n * w; idx += blockDim.x) M[idx] = 0.0;
    __syncthreads();
    if (j < n) {
        M[j * w + j] = diag[j];
        M[j * w + n + j] = 1.0;
    }
    __syncthreads();
    if (j == 0) {   // face order (faces may repeat a cell pair on a fine level)
        for (int f = 0; f < m; ++f) {
            M[la[f] * w + ua[f]] = M[la[f] * w + ua[f]] + upper[f];
            M[ua[f] * w + la[f]] = M[ua[f] * w + la[f]] + lower[f];
        }
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double piv = M[k * w + k];
        __syncthreads();
        if (j < w) M[k * w + j] = M[k * w + j] / piv;
        __syncthreads();
        if (j < n) fac[j] = M[j * w + k];
        __syncthreads();
        if (j < w) {
            const double pk = M[k * w + j];
            for (int i = 0; i < n; ++i)
                if (i != k) M[i * w + j] = fma(-fac[i], pk, M[i * w + j]);
        }
        __syncthreads();
    }
    if (j < n)
        for (int i = 0; i < n; ++i) ainv[i * n + j] = M[i * w + n + j];
}

// ---------------------------------------------------------------------------
// device: V-cycle pieces (every kernel is a no-op once the solve is done),
// for T = double (GAMG) and T = float (GAMG32: the same steps in FP32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double fmaT(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return fmaf(a, b, c); }

template <typename T>
__global__ __launch_bounds__(kWG) void k_vc_pre(int N, const T* __restrict__ sm,
                                                const T* __restrict__ r, T* __restrict__ z,
                                                const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    T sv[4], rv[4], zv[4];
    load4(sm, t.i0, N, sv);
    load4(r, t.i0, N, rv);
#pragma unroll
    for (int k = 0; k < 4; ++k) zv[k] = sv[k] * rv[k];
    store4(z, t.i0, N, zv);
}

// GAMG32 entry on the finest level: r32 = (float) r, z32 = sm32 .* r32
__global__ __launch_bounds__(kWG) void k_vc_narrow_pre(int N, const float* __restrict__ sm,
                                                       const double* __restrict__ r,
                                                       float* __restrict__ r32,
                                                       float* __restrict__ z32,
                                                       const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    double rv[4];
    float sv[4], fv[4], zv[4];
    load4(sm, t.i0, N, sv);
    load4(r, t.i0, N, rv);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        fv[k] = (float)rv[k];
        zv[k] = sv[k] * fv[k];
    }
    store4(r32, t.i0, N, fv);
    store4(z32, t.i0, N, zv);
}

// One damped-Jacobi sweep zn = fma(sm, r - A z, z).  RED (finest level, FP64):
// rho = (r, zn), beta = rho / rho_prev.
template <class P, bool RED, typename T>
__global__ __launch_bounds__(kWG) void k_vc_sweep(Mat<T> m, int N, const T* __restrict__ sm,
                                                  const T* __restrict__ r,
                                                  const T* __restrict__ z, T* __restrict__ zn,
                                                  KState* st, RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<T> sh;
    __shared__ T sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return z[c]; }, sh, sh_y);
    T sv[4], rv[4], zv[4];
    load4(sm, t.i0, N, sv);
    load4(r, t.i0, N, rv);
    load4(z, t.i0, N, zv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        zv[k] = fmaT(sv[k], rv[k] - sh_y[threadIdx.x * kLeaf + k], zv[k]);
        if (RED && t.i0 + k < N) leaf[0] = fma((double)rv[k], (double)zv[k], leaf[0]);
    }
    store4nt(zn, t.i0, N, zv);
    if constexpr (RED) {
        finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
            st->rho1 = st->rho0;
            st->rho0 = v[0];
            st->beta = st->rho0 / st->rho1;
        });
    }
}

// GAMG32's last sweep on the finest level: zn = fmaf(sm, (float) r - A z, z),
// out = (double) zn, rho = (r, out) in FP64 and the flexible direction
// factor beta = (rho - gamma) / rho_prev, gamma = (r, z_prev) from k_pcg2.
template <class P>
__global__ __launch_bounds__(kWG) void k_vc_sweep_out(Mat<float> m, int N,
                                                      const float* __restrict__ sm,
                                                      const double* __restrict__ r,
                                                      const float* __restrict__ z,
                                                      double* __restrict__ out, KState* st,
                                                      RedWs red) {
    if (st->done) return;
    __shared__ typename P::template Shared<float> sh;
    __shared__ float sh_y[kTile];
    __shared__ RedShared shr;
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return z[c]; }, sh, sh_y);
    float sv[4], zv[4];
    double rv[4], ov[4];
    load4(sm, t.i0, N, sv);
    load4(r, t.i0, N, rv);
    load4(z, t.i0, N, zv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        zv[k] = fmaf(sv[k], (float)rv[k] - sh_y[threadIdx.x * kLeaf + k], zv[k]);
        ov[k] = (double)zv[k];
        if (t.i0 + k < N) leaf[0] = fma(rv[k], ov[k], leaf[0]);
    }
    store4nt(out, t.i0, N, ov);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = (st->rho0 - st->aux[kGammaSlot]) / st->rho1;
    });
}

// t = r - A z
template <class P, typename T>
__global__ __launch_bounds__(kWG) void k_vc_resid(Mat<T> m, int N, const T* __restrict__ r,
                                                  const T* __restrict__ z, T* __restrict__ tr,
                                                  const KState* st) {
    if (st->done) return;
    __shared__ typename P::template Shared<T> sh;
    __shared__ T sh_y[kTile];
    TileRows t(N);
    P::run(m, N, t.row0, t.nrows, [&](int c) { return z[c]; }, sh, sh_y);
    T rv[4], tv[4];
    load4(r, t.i0, N, rv);
#pragma unroll
    for (int k = 0; k < 4; ++k) tv[k] = rv[k] - sh_y[threadIdx.x * kLeaf + k];
    store4nt(tr, t.i0, N, tv);
}

// Coarse levels (below the finest): aggregation scatters a row's columns over
// thousands of rows, so offset tables rarely fit and these kernels read the
// int32-column slices, which every slice has.  A level of a few thousand to a
// few hundred thousand rows fills only part of the chip, so the per-tile
// kernels above wait on memory latency: each wave walks 4 slices one after
// another, 8 slots at a time.  Here each wave takes one 64-row slice (one row
// per lane), with kCoarseBatch slots' loads in flight before the gathers.
// Same products in the same order as every other schedule: the results are
// bit-identical.
constexpr int kCoarseBatch = 32;
constexpr int kSlicesPerWG = kWG / kSlice;

__device__ __forceinline__ int wave_slice() {
    return (int)blockIdx.x * kSlicesPerWG + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_vc_sweep_w(Mat<T> m, int N, const T* __restrict__ sm,
                                                    const T* __restrict__ r,
                                                    const T* __restrict__ z, T* __restrict__ zn,
                                                    const KState* st) {
    if (st->done) return;
    const int s = wave_slice();
    if (s * kSlice >= N) return;   // whole waves
    const int lane = threadIdx.x & 63;
    const int a = m.soff[s];
    const T y = sell_row<T, kCoarseBatch>(m.scol, m.sval, a, (m.soff[s + 1] - a) / kSlice, lane,
                                          [&](int c) { return z[c]; });
    const int row = s * kSlice + lane;
    if (row < N) zn[row] = fmaT(sm[row], r[row] - y, z[row]);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_vc_resid_w(Mat<T> m, int N, const T* __restrict__ r,
                                                    const T* __restrict__ z, T* __restrict__ tr,
                                                    const KState* st) {
    if (st->done) return;
    const int s = wave_slice();
    if (s * kSlice >= N) return;
    const int lane = threadIdx.x & 63;
    const int a = m.soff[s];
    const T y = sell_row<T, kCoarseBatch>(m.scol, m.sval, a, (m.soff[s + 1] - a) / kSlice, lane,
                                          [&](int c) { return z[c]; });
    const int row = s * kSlice + lane;
    if (row < N) tr[row] = r[row] - y;
}

// rc[C] = sum of t over C's cells, ascending, and the next level's first
// sweep from zero, zc0 = smc .* rc.  Members are read 4 at a time (their
// loads in flight together), summed one by one in order.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_vc_restrict(int nc, const int* __restrict__ mptr,
                                                      const int* __restrict__ mlist,
                                                      const T* __restrict__ t,
                                                      const T* __restrict__ smc,
                                                      T* __restrict__ rc, T* __restrict__ zc0,
                                                      const KState* st) {
    if (st->done) return;
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= nc) return;
    const int p0 = mptr[c], p1 = mptr[c + 1];
    T s = T(0);
    for (int p = p0; p < p1; p += 4) {
        int q[4];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = mlist[min(p + u, p1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = t[q[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (p + u < p1) s = s + v[u];
    }
    rc[c] = s;
    zc0[c] = smc[c] * s;
}

// z = fma(cscale, zc[agg], z)   (a separate pass: folding it into the next
// sweep's gathers costs two dependent loads per nonzero, measured slower)
template <typename T>
__global__ __launch_bounds__(kWG) void k_vc_prolong(int N, T cscale, const int* __restrict__ agg,
                                                    const T* __restrict__ zc,
                                                    T* __restrict__ z, const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    T zv[4];
    load4(z, t.i0, N, zv);
    int a[4];
    if (t.i0 + 3 < N) {   // the row's 4 aggregates in one 16-byte load
        const int4 q = *reinterpret_cast<const int4*>(agg + t.i0);
        a[0] = q.x;
        a[1] = q.y;
        a[2] = q.z;
        a[3] = q.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = t.i0 + k < N ? agg[t.i0 + k] : 0;
    }
    T c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = zc[a[k]];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) zv[k] = fmaT(cscale, c[k], zv[k]);
    store4(z, t.i0, N, zv);
}

// The levels from the first one of at most kTail cells down to the last, as
// one workgroup: the same steps and arithmetic as the per-level kernels (CSR
// rows summed in storage order), separated by workgroup barriers.
template <typename T>
struct TailLevel {
    int n;
    const int* rowptr;
    const int* col;
    const T* val;
    const T* sm;
    const int* agg;     // to the next tail level
    const int* mptr;
    const int* mlist;
    T* r;
    T* z;
    T* z2;
    T* t;
};

constexpr int kTail = 256;        // one workgroup runs a level in ~10 us per step: only tiny levels
constexpr int kTailThreads = 1024;

// Row i of a tail level: a batch's column and value loads, then its gathers,
// are in flight together; the products are summed one by one in storage order.
template <typename T>
__device__ __forceinline__ T csr_row(const TailLevel<T>& L, int i, const T* x) {
    const int a = L.rowptr[i], b = L.rowptr[i + 1];
    T s = T(0);
    for (int j = a; j < b; j += 8) {
        int c[8];
        T v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = min(j + u, b - 1);
            c[u] = L.col[q];
            v[u] = L.val[q];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j + u < b) s = s + v[u] * xv[u];
    }
    return s;
}

// On entry lv[0].r holds the rhs and lv[0].z its first sweep from zero; on
// exit lv[0].z holds the cycle's result.
template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_vc_tail(const TailLevel<T>* __restrict__ lv,
                                                          int nl, int nu, T cscale,
                                                          const T* __restrict__ ainv,
                                                          int dense_n, const KState* st) {
    if (st->done) return;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    T* cur[kAmgLevels];
    for (int l = 0; l + 1 < nl; ++l) {
        const TailLevel<T>& L = lv[l];
        T* z = L.z;
        T* z2 = L.z2;
        for (int s = 1; s < nu; ++s) {
            for (int i = tid; i < L.n; i += nt)
                z2[i] = fmaT(L.sm[i], L.r[i] - csr_row(L, i, z), z[i]);
            __syncthreads();
            T* tmp = z;
            z = z2;
            z2 = tmp;
        }
        for (int i = tid; i < L.n; i += nt) L.t[i] = L.r[i] - csr_row(L, i, z);
        __syncthreads();
        const TailLevel<T>& C = lv[l + 1];
        for (int c = tid; c < C.n; c += nt) {
            const int p0 = L.mptr[c], p1 = L.mptr[c + 1];
            T s = T(0);
            for (int p = p0; p < p1; p += 4) {
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = L.t[L.mlist[min(p + u, p1 - 1)]];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (p + u < p1) s = s + v[u];
            }
            C.r[c] = s;
            C.z[c] = C.sm[c] * s;
        }
        __syncthreads();
        cur[l] = z;
    }
    const TailLevel<T>& E = lv[nl - 1];
    if (dense_n) {
        for (int i = tid; i < E.n; i += nt) {
            T s = T(0);
            for (int j = 0; j < E.n; ++j) s = s + ainv[i * E.n + j] * E.r[j];
            E.z[i] = s;
        }
        __syncthreads();
    }
    const T* zc = E.z;
    for (int l = nl - 2; l >= 0; --l) {
        const TailLevel<T>& L = lv[l];
        T* z = cur[l];
        T* z2 = (z == L.z) ? L.z2 : L.z;
        for (int i = tid; i < L.n; i += nt) z[i] = fmaT(cscale, zc[L.agg[i]], z[i]);
        __syncthreads();
        for (int s = 0; s < nu; ++s) {
            for (int i = tid; i < L.n; i += nt)
                z2[i] = fmaT(L.sm[i], L.r[i] - csr_row(L, i, z), z[i]);
            __syncthreads();
            T* tmp = z;
            z = z2;
            z2 = tmp;
        }
        if (z != L.z) {
            for (int i = tid; i < L.n; i += nt) L.z[i] = z[i];
            __syncthreads();
        }
        zc = L.z;
    }
}

// z = Ainv r on the last level (n <= kDense)
template <typename T>
__global__ __launch_bounds__(kDense) void k_vc_dense(int n, const T* __restrict__ ainv,
                                                     const T* __restrict__ r, T* __restrict__ z,
                                                     const KState* st) {
    if (st->done) return;
    const int i = threadIdx.x;
    if (i >= n) return;
    T s = T(0);
    for (int j = 0; j < n; ++j) s = s + ainv[i * n + j] * r[j];
    z[i] = s;
}

__global__ __launch_bounds__(kBlk) void k_narrow(long long n, const double* __restrict__ a,
                                                 float* __restrict__ b) {
    const long long i = (long long)blockIdx.x * kBlk + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// rho = (r, z), beta = rho / rho_prev (after a one-level V-cycle)
__global__ __launch_bounds__(kWG) void k_pcg_rho(int N, const double* __restrict__ r,
                                                 const double* __restrict__ z, KState* st,
                                                 RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
    double rv[4], zv[4];
    load4(r, t.i0, N, rv);
    load4(z, t.i0, N, zv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = fma(rv[k], zv[k], leaf[0]);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = st->rho0 / st->rho1;
    });
}

// one-level GAMG32: out = (double) z32, rho = (r, out), flexible beta as k_vc_sweep_out
__global__ __launch_bounds__(kWG) void k_pcg_rho32(int N, const double* __restrict__ r,
                                                   const float* __restrict__ z32,
                                                   double* __restrict__ out, KState* st,
                                                   RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
    double rv[4], ov[4];
    float zv[4];
    load4(r, t.i0, N, rv);
    load4(z32, t.i0, N, zv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ov[k] = (double)zv[k];
        if (t.i0 + k < N) leaf[0] = fma(rv[k], ov[k], leaf[0]);
    }
    store4nt(out, t.i0, N, ov);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        st->rho1 = st->rho0;
        st->rho0 = v[0];
        st->beta = (st->rho0 - st->aux[kGammaSlot]) / st->rho1;
    });
}

// z = dinv r and rho = (r, z) (initial diagonal preconditioning)
__global__ __launch_bounds__(kWG) void k_pcg_diag0(int N, const double* __restrict__ dinv,
                                                   const double* __restrict__ r,
                                                   double* __restrict__ z, KState* st,
                                                   RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    TileRows t(N);
    double dv[4], rv[4], zv[4];
    load4(dinv, t.i0, N, dv);
    load4(r, t.i0, N, rv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        zv[k] = dv[k] * rv[k];
        if (t.i0 + k < N) leaf[0] = fma(rv[k], zv[k], leaf[0]);
    }
    store4(z, t.i0, N, zv);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) { st->rho0 = v[0]; });
}

// p = z (k = 1) or fma(beta, p_prev, z); x = fma(alpha_prev, p_prev, x) (deferred)
__global__ __launch_bounds__(kWG) void k_pcg_dir(int N, const double* __restrict__ z,
                                                 double* __restrict__ pb0,
                                                 double* __restrict__ pb1,
                                                 double* __restrict__ x, const KState* st) {
    if (st->done) return;
    StreamRows t(N);
    const int k = st->k + 1;
    const double* p_old = (k & 1) ? pb0 : pb1;
    double* p_new = (k & 1) ? pb1 : pb0;
    double zv[4], pv[4], xv[4];
    load4(z, t.i0, N, zv);
    if (k == 1) {
        store4(p_new, t.i0, N, zv);
        return;
    }
    const double beta = st->beta;
    const double alpha_prev = st->alpha;
    load4(p_old, t.i0, N, pv);
    load4(x, t.i0, N, xv);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        xv[kk] = fma(alpha_prev, pv[kk], xv[kk]);
        pv[kk] = fma(beta, pv[kk], zv[kk]);
    }
    store4(p_new, t.i0, N, pv);
    store4(x, t.i0, N, xv);
}

// r = fma(-alpha, q, r); ||r||, history, stop test.  PC 0: rho = (r, r);
// PC 1: z = dinv r, rho = (r, z); PC 2, 3: rho is left to the V-cycle;
// PC 4: rho is left to the DIC sweeps (k_pcg_rz).
// PC 2 also writes the V-cycle's first sweep from zero, z = sm .* r (dinv = sm);
// PC 3 writes r32 = (float) r and z32 = sm32 .* r32.
template <int C, int PC>
__global__ __launch_bounds__(kWG) void k_pcg2(int N, double* __restrict__ r,
                                              const double* __restrict__ q,
                                              const double* __restrict__ dinv,
                                              double* __restrict__ z,
                                              const float* __restrict__ sm32,
                                              float* __restrict__ r32, float* __restrict__ z32,
                                              KState* st, RedWs red) {
    if (st->done) return;
    constexpr int NV = (PC == 1 || PC == 3) ? 2 : 1;
    __shared__ RedShared shr;
    const SuperTile<C> sup(N, true);
    const double alpha = st->alpha;
    double leaf[C][NV];
    double rv[C][4], zv[C][4];
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int v = 0; v < NV; ++v) leaf[c][v] = 0.0;
        if (c >= sup.nt) continue;
        const TileRows t(N, sup.tile0 + c);
        double qv[4], dv[4];
        load4(r, t.i0, N, rv[c]);
        load4(q, t.i0, N, qv);
        if (PC == 1 || PC == 2) load4(dinv, t.i0, N, dv);
        if (PC == 3) load4(z, t.i0, N, zv[c]);   // the previous preconditioned residual
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            rv[c][kk] = fma(-alpha, qv[kk], rv[c][kk]);
            if (PC == 1 || PC == 2) zv[c][kk] = dv[kk] * rv[c][kk];
            if (t.i0 + kk < N) {
                leaf[c][0] = fma(rv[c][kk], rv[c][kk], leaf[c][0]);
                if (PC == 1 || PC == 3)
                    leaf[c][NV - 1] = fma(rv[c][kk], zv[c][kk], leaf[c][NV - 1]);
            }
        }
    }
    SlotReduce<NV, C> tr;
    tr.begin(red, sup.tile0, sup.nt, leaf, shr);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (c >= sup.nt) continue;
        const int i0 = TileRows(N, sup.tile0 + c).i0;
        store4(r, i0, N, rv[c]);
        if (PC == 1 || PC == 2) store4(z, i0, N, zv[c]);
        if (PC == 3) {
            float sv[4], fv[4], gv[4];
            load4(sm32, i0, N, sv);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                fv[kk] = (float)rv[c][kk];
                gv[kk] = sv[kk] * fv[kk];
            }
            store4(r32, i0, N, fv);
            store4(z32, i0, N, gv);
        }
    }
    tr.end(red, st, shr, [&](const double (&v)[NV]) {
        const double norm = sqrt(v[0]);
        record_norm(st, norm);
        if (PC < 2) {
            st->rho1 = st->rho0;
            st->rho0 = v[NV - 1];
            st->beta = st->rho0 / st->rho1;
        }
        if (PC == 3) st->aux[kGammaSlot] = v[1];
        if (stop_after(st, !(norm > st->error))) st->done = 1;
    });
}

// ---------------------------------------------------------------------------
// host: hierarchy on the device
// ---------------------------------------------------------------------------
// V: a vector of trivially copyable elements (std::vector or agglom::hvec)
template <typename V>
int upload_vec(gis_context* h, DevMem& d, const V& v) {
    using T = typename V::value_type;
    GIS_HIP(h, d.ensure(sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty())
        GIS_HIP(h, hipMemcpyAsync(d.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice,
                                  h->stream));
    return kOK;
}

template <typename V>
int download(gis_context* h, V& out, const DevMem& d, size_t n) {
    using T = typename V::value_type;
    out.resize(n);
    if (n)
        GIS_HIP(h, hipMemcpyAsync(out.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost,
                                  h->stream));
    return kOK;
}

gis_context* level_ctx(gis_context* h, int l) { return l == 0 ? h : h->amg->lv[l].ctx; }

int amg_build(gis_context* h) {
    using agglom::HostLevel;
    AmgData& A = *h->amg;
    A.clear();
    HostLevel fine;
    fine.n = h->N;
    fine.m = h->M;
    GIS_TRY(download(h, fine.la, h->lower_addr, (size_t)h->M));
    GIS_TRY(download(h, fine.ua, h->upper_addr, (size_t)h->M));
    GIS_TRY(download(h, fine.lower, h->lower, (size_t)h->M));
    GIS_TRY(download(h, fine.upper, h->upper, (size_t)h->M));
    GIS_TRY(download(h, fine.diag, h->diag, (size_t)h->N));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    A.lv.emplace_back();
    A.lv[0].n = fine.n;
    A.lv[0].m = fine.m;
    HostLevel cur = std::move(fine);
    while (cur.n > kDense && (int)A.lv.size() < kAmgLevels) {
        HostLevel next;
        agglom::Transfer t;
        if (!agglom::coarsen(cur, next, t)) break;   // stagnation
        // device data of the transfer cur -> next
        AmgLevel& L = A.lv.back();
        L.nc = next.n;
        GIS_TRY(upload_vec(h, L.agg, t.agg));
        GIS_TRY(upload_vec(h, L.mptr, t.mptr));
        GIS_TRY(upload_vec(h, L.mlist, t.mlist));
        GIS_TRY(upload_vec(h, L.iptr, t.iptr));
        GIS_TRY(upload_vec(h, L.ilist, t.ilist));
        GIS_TRY(upload_vec(h, L.fptr, t.fptr));
        GIS_TRY(upload_vec(h, L.flist, t.flist));
        // the next level as a pattern of its own
        A.lv.emplace_back();
        AmgLevel& C = A.lv.back();
        C.n = next.n;
        C.m = next.m;
        C.ctx = new (std::nothrow) gis_context();
        if (!C.ctx) return fail(h, kErrNoMem, "GAMG level context");
        C.ctx->device = h->device;
        C.ctx->stream = h->stream;
        C.ctx->own_stream = false;
        C.ctx->sym_enable = h->sym_enable;
        GIS_TRY(upload_vec(h, C.ctx->lower_addr, next.la));
        GIS_TRY(upload_vec(h, C.ctx->upper_addr, next.ua));
        // the uploads read t and next, which go out of scope below
        GIS_HIP(h, hipStreamSynchronize(h->stream));
        if (build_pattern(C.ctx, next.n, next.m, C.ctx->lower_addr.as<int>(),
                          C.ctx->upper_addr.as<int>()) != kOK)
            return fail(h, kErrHip, "GAMG level pattern: " + C.ctx->err);
        cur = std::move(next);
    }
    A.dense_n = cur.n <= kDense ? cur.n : 0;
    for (size_t l = 0; l < A.lv.size(); ++l) {
        AmgLevel& L = A.lv[l];
        const size_t n = (size_t)L.n;
        GIS_HIP(h, L.sm.ensure(sizeof(double) * n));
        GIS_HIP(h, L.vecs.ensure(sizeof(double) * 4 * n));
    }
    if (A.dense_n) {
        GIS_HIP(h, A.ainv.ensure(sizeof(double) * A.dense_n * A.dense_n));
        GIS_HIP(h, A.dense_work.ensure(sizeof(double) * 2 * A.dense_n * A.dense_n));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    A.built = true;
    A.values_ok = false;
    return kOK;
}

int amg_values(gis_context* h) {
    AmgData& A = *h->amg;
    for (size_t l = 0; l < A.lv.size(); ++l) {
        AmgLevel& L = A.lv[l];
        gis_context* c = level_ctx(h, (int)l);
        if (l > 0) GIS_TRY(gather_values(c));
        hipLaunchKernelGGL(k_amg_sm, grid_of(L.n), dim3(kBlk), 0, h->stream, L.n, A.omega,
                           c->diag.as<double>(), L.sm.as<double>());
        GIS_HIP(h, hipGetLastError());
        if (l + 1 < A.lv.size()) {
            gis_context* n = A.lv[l + 1].ctx;
            const AmgLevel& C = A.lv[l + 1];
            GIS_HIP(h, n->diag.ensure(sizeof(double) * std::max(C.n, 1)));
            GIS_HIP(h, n->lower.ensure(sizeof(double) * std::max(C.m, 1)));
            GIS_HIP(h, n->upper.ensure(sizeof(double) * std::max(C.m, 1)));
            hipLaunchKernelGGL(k_amg_cdiag, grid_of(C.n), dim3(kBlk), 0, h->stream, C.n,
                               L.mptr.as<int>(), L.mlist.as<int>(), L.iptr.as<int>(),
                               L.ilist.as<int>(), c->diag.as<double>(), c->lower.as<double>(),
                               c->upper.as<double>(), n->diag.as<double>());
            GIS_HIP(h, hipGetLastError());
            if (C.m > 0) {
                hipLaunchKernelGGL(k_amg_cface, grid_of(C.m), dim3(kBlk), 0, h->stream, C.m,
                                   L.fptr.as<int>(), L.flist.as<int>(), c->lower.as<double>(),
                                   c->upper.as<double>(), n->lower.as<double>(),
                                   n->upper.as<double>());
                GIS_HIP(h, hipGetLastError());
            }
        }
    }
    if (A.dense_n) {
        gis_context* c = level_ctx(h, (int)A.lv.size() - 1);
        hipLaunchKernelGGL(k_amg_dense_inv, dim3(1), dim3(2 * kDense), 0, h->stream, A.dense_n,
                           c->M, c->lower_addr.as<int>(), c->upper_addr.as<int>(),
                           c->diag.as<double>(), c->lower.as<double>(), c->upper.as<double>(),
                           A.dense_work.as<double>(), A.ainv.as<double>());
        GIS_HIP(h, hipGetLastError());
    }
    // the coarse tail (levels of at most kTail cells below the finest) as one workgroup
    A.tail_start = 0;
    A.tail_levels = 0;
    for (size_t l = 1; l < A.lv.size(); ++l)
        if (A.lv[l].n <= kTail) {
            A.tail_start = (int)l;
            break;
        }
    if (A.tail_start) {
        std::vector<TailLevel<double>> meta;
        for (size_t l = A.tail_start; l < A.lv.size(); ++l) {
            AmgLevel& L = A.lv[l];
            gis_context* c = level_ctx(h, (int)l);
            double* v = L.vecs.as<double>();
            const size_t n = (size_t)L.n;
            const bool last = l + 1 == A.lv.size();
            meta.push_back(TailLevel<double>{L.n, c->rowptr.as<int>(), c->col.as<int>(),
                                     c->val.as<double>(), L.sm.as<double>(),
                                     last ? nullptr : L.agg.as<int>(),
                                     last ? nullptr : L.mptr.as<int>(),
                                     last ? nullptr : L.mlist.as<int>(), v + 3 * n, v, v + n,
                                     v + 2 * n});
        }
        A.tail_levels = (int)meta.size();
        GIS_TRY(upload_vec(h, A.tail, meta));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    A.values_ok = true;
    A.values32_ok = false;
    return kOK;
}

// FP32 copies for GAMG32: level values, sm32 = (float) sm, (float) ainv
int amg_values32(gis_context* h) {
    AmgData& A = *h->amg;
    for (size_t l = 0; l < A.lv.size(); ++l) {
        AmgLevel& L = A.lv[l];
        GIS_TRY(ensure_val32(level_ctx(h, (int)l)));
        GIS_HIP(h, L.sm32.ensure(sizeof(float) * std::max(L.n, 1)));
        GIS_HIP(h, L.vecs32.ensure(sizeof(float) * 4 * std::max(L.n, 1)));
        hipLaunchKernelGGL(k_narrow, grid_of(L.n), dim3(kBlk), 0, h->stream, (long long)L.n,
                           L.sm.as<double>(), L.sm32.as<float>());
        GIS_HIP(h, hipGetLastError());
    }
    if (A.dense_n) {
        const long long nn = (long long)A.dense_n * A.dense_n;
        GIS_HIP(h, A.ainv32.ensure(sizeof(float) * nn));
        hipLaunchKernelGGL(k_narrow, grid_of(nn), dim3(kBlk), 0, h->stream, nn,
                           A.ainv.as<double>(), A.ainv32.as<float>());
        GIS_HIP(h, hipGetLastError());
    }
    if (A.tail_start) {
        std::vector<TailLevel<float>> meta;
        for (size_t l = A.tail_start; l < A.lv.size(); ++l) {
            AmgLevel& L = A.lv[l];
            gis_context* c = level_ctx(h, (int)l);
            float* v = L.vecs32.as<float>();
            const size_t n = (size_t)L.n;
            const bool last = l + 1 == A.lv.size();
            meta.push_back(TailLevel<float>{L.n, c->rowptr.as<int>(), c->col.as<int>(),
                                            c->val32.as<float>(), L.sm32.as<float>(),
                                            last ? nullptr : L.agg.as<int>(),
                                            last ? nullptr : L.mptr.as<int>(),
                                            last ? nullptr : L.mlist.as<int>(), v + 3 * n, v,
                                            v + n, v + 2 * n});
        }
        GIS_TRY(upload_vec(h, A.tail32, meta));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    A.values32_ok = true;
    return kOK;
}

int amg_prepare(gis_context* h) {
    if (h->n_cpl > 0)
        return fail(h, kErrState, "GAMG does not take interface entries (gis_set_interfaces)");
    if (h->batch.ready) return fail(h, kErrState, "GAMG needs a single-system pattern");
    if (!h->amg) {
        h->amg = new (std::nothrow) AmgData();
        if (!h->amg) return fail(h, kErrNoMem, "GAMG data");
    }
    if (!h->amg->built) {
        // the host set-up allocates and starts threads: no C++ exception may
        // cross the C ABI
        int st = kOK;
        try {
            st = amg_build(h);
        } catch (const std::bad_alloc&) {
            st = fail(h, kErrNoMem, "GAMG set-up: host memory exhausted");
        } catch (const std::exception& e) {
            st = fail(h, kErrNoMem, std::string("GAMG set-up: ") + e.what());
        }
        if (st != kOK) {
            h->amg->clear();
            return st;
        }
    }
    if (!h->amg->values_ok) GIS_TRY(amg_values(h));
    return kOK;
}

template <typename T>
struct LevelBufs {
    static T* vecs(AmgLevel& L);
    static T* sm(AmgLevel& L);
};
template <>
double* LevelBufs<double>::vecs(AmgLevel& L) { return L.vecs.as<double>(); }
template <>
double* LevelBufs<double>::sm(AmgLevel& L) { return L.sm.as<double>(); }
template <>
float* LevelBufs<float>::vecs(AmgLevel& L) { return L.vecs32.as<float>(); }
template <>
float* LevelBufs<float>::sm(AmgLevel& L) { return L.sm32.as<float>(); }

template <typename T>
Mat<T> level_mat(const gis_context* c);
template <>
Mat<double> level_mat<double>(const gis_context* c) { return mat64(c); }
template <>
Mat<float> level_mat<float>(const gis_context* c) { return mat32(c); }

// V-cycle of level l on rhs r (type T); returns the buffer holding z.  On
// entry the level's z holds the first sweep from zero, sm .* r (written by
// whoever formed r).  On the finest level the last sweep goes to `out` (FP64)
// together with rho = (r64, out); r64 is the FP64 residual (T = float).
template <class P, typename T>
const T* vcycle(gis_context* h, int l, const T* r, double* out, const double* r64) {
    AmgData& A = *h->amg;
    AmgLevel& L = A.lv[l];
    gis_context* c = level_ctx(h, l);
    const int N = L.n;
    T* za = LevelBufs<T>::vecs(L);
    T* zb = za + N;
    T* t = za + 2 * (size_t)N;
    const dim3 g = tiles_grid(N);
    const Mat<T> m = level_mat<T>(c);
    const T* sm = LevelBufs<T>::sm(L);
    const T* ainv = nullptr;
    if constexpr (std::is_same<T, double>::value) ainv = A.ainv.as<double>();
    else ainv = A.ainv32.as<float>();
    if (l + 1 == (int)A.lv.size()) {
        T* dst = za;
        if constexpr (std::is_same<T, double>::value) {
            if (l == 0) dst = out;   // a one-level hierarchy writes z straight into out
        }
        if (A.dense_n)
            hipLaunchKernelGGL((k_vc_dense<T>), dim3(1), dim3(kDense), 0, h->stream, N, ainv, r,
                               dst, h->st);
        else
            hipLaunchKernelGGL((k_vc_pre<T>), g, dim3(kWG), 0, h->stream, N, sm, r, dst, h->st);
        if (l == 0) {   // rho = (r, z) for a one-level hierarchy
            if constexpr (std::is_same<T, double>::value)
                hipLaunchKernelGGL(k_pcg_rho, g, dim3(kWG), 0, h->stream, N, r64, out, h->st,
                                   h->red);
            else
                hipLaunchKernelGGL(k_pcg_rho32, g, dim3(kWG), 0, h->stream, N, r64, za, out,
                                   h->st, h->red);
        }
        return dst;
    }
    // non-reducing sweeps and the residual: per-slice waves below the finest level
    const bool wave = l > 0;
    const dim3 gw((unsigned)((N + kSlicesPerWG * kSlice - 1) / (kSlicesPerWG * kSlice)));
    auto sweep = [&](const T* zin, T* zout) {
        if (wave)
            hipLaunchKernelGGL((k_vc_sweep_w<T>), gw, dim3(kWG), 0, h->stream, m, N, sm, r, zin,
                               zout, h->st);
        else
            hipLaunchKernelGGL((k_vc_sweep<P, false, T>), g, dim3(kWG), 0, h->stream, m, N, sm,
                               r, zin, zout, h->st, h->red);
    };
    for (int s = 1; s < A.nu; ++s) {
        sweep(za, zb);
        std::swap(za, zb);
    }
    if (wave)
        hipLaunchKernelGGL((k_vc_resid_w<T>), gw, dim3(kWG), 0, h->stream, m, N, r, za, t, h->st);
    else
        hipLaunchKernelGGL((k_vc_resid<P, T>), g, dim3(kWG), 0, h->stream, m, N, r, za, t, h->st);
    AmgLevel& C = A.lv[l + 1];
    T* zc0 = LevelBufs<T>::vecs(C);
    T* rc = zc0 + 3 * (size_t)C.n;
    hipLaunchKernelGGL((k_vc_restrict<T>), grid_of(C.n), dim3(kBlk), 0, h->stream, C.n,
                       L.mptr.as<int>(), L.mlist.as<int>(), t, LevelBufs<T>::sm(C), rc, zc0,
                       h->st);
    const T* zc;
    if (l + 1 == A.tail_start) {
        const TailLevel<T>* meta;
        if constexpr (std::is_same<T, double>::value) meta = A.tail.as<TailLevel<double>>();
        else meta = A.tail32.as<TailLevel<float>>();
        hipLaunchKernelGGL((k_vc_tail<T>), dim3(1), dim3(kTailThreads), 0, h->stream, meta,
                           A.tail_levels, A.nu, (T)A.cscale, A.dense_n ? ainv : nullptr,
                           A.dense_n, h->st);
        zc = zc0;
    } else {
        zc = vcycle<P, T>(h, l + 1, rc, nullptr, nullptr);
    }
    hipLaunchKernelGGL((k_vc_prolong<T>), g, dim3(kWG), 0, h->stream, N, (T)A.cscale,
                       L.agg.as<int>(), zc, za, h->st);
    for (int s = 0; s < A.nu; ++s) {
        const bool last_fine = (l == 0 && s + 1 == A.nu);
        if (last_fine) {
            if constexpr (std::is_same<T, double>::value)
                hipLaunchKernelGGL((k_vc_sweep<P, true, double>), g, dim3(kWG), 0, h->stream, m,
                                   N, sm, r, za, out, h->st, h->red);
            else
                hipLaunchKernelGGL((k_vc_sweep_out<P>), g, dim3(kWG), 0, h->stream, m, N, sm,
                                   r64, za, out, h->st, h->red);
            return za;
        }
        sweep(za, zb);
        std::swap(za, zb);
    }
    return za;
}

}  // namespace

AmgData::~AmgData() { clear(); }

void AmgData::clear() {
    for (auto& L : lv) delete L.ctx;
    lv.clear();
    ainv.release();
    ainv32.release();
    dense_work.release();
    tail.release();
    tail32.release();
    dense_n = 0;
    tail_start = tail_levels = 0;
    built = false;
    values_ok = false;
    values32_ok = false;
    dic_levels = false;
    dic_values = false;
    dic_kind = 0;
    dic_ptr_f.clear();
    dic_ptr_b.clear();
    dic_cells_f.release();
    dic_cells_b.release();
    dic_rd.release();
}

// ---------------------------------------------------------------------------
// DIC, OpenFOAM's DICPreconditioner restated from its published algorithm (the
// reference has no preconditioners, README.md:71-73):
//   rD = diag; for each face f in face order: rD[u] -= upper[f] upper[f] / rD[l];
//   rD = 1 / rD;
//   z = rD .* r; for f in face order: z[u] -= rD[u] upper[f] z[l];
//   for f in reverse face order: z[l] -= rD[l] upper[f] z[u].
// The face loops are sequential; here cells run in dependency levels (a cell's
// forward level is one more than its lower neighbours', its backward level one
// more than its upper neighbours'), one launch per level, and each cell applies
// its own faces in the order the face loop would: its lower faces in face
// order (the CSR row's lower entries, perm codes f), its upper faces in reverse
// face order.  Every value is then formed by the same operations in the same
// order as the face loops form it.
//
// DILU (OpenFOAM's DILUPreconditioner, for asymmetric matrices) runs the same
// kernels with other coefficient arrays: rD[u] -= upper[f] lower[f] / rD[l];
// the forward sweep reads lower, the backward one upper; M^-T exchanges them.
// With lower == upper its bits are DIC's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlk) void k_dic_factor(int ncell, const int* __restrict__ cells,
                                                     const int* __restrict__ rowptr,
                                                     const int* __restrict__ nlow,
                                                     const int* __restrict__ col,
                                                     const int* __restrict__ perm,
                                                     const double* __restrict__ diag,
                                                     const double* __restrict__ upper,
                                                     const double* __restrict__ lower,
                                                     double* __restrict__ F) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= ncell) return;
    const int c = cells[i];
    double f = diag[c];
    const int a = rowptr[c];
    const int e = a + nlow[c];
    for (int p = a; p < e; ++p) f = f - upper[perm[p]] * lower[perm[p]] / F[col[p]];
    F[c] = f;
}

__global__ __launch_bounds__(kBlk) void k_dic_fwd(int ncell, const int* __restrict__ cells,
                                                  const int* __restrict__ rowptr,
                                                  const int* __restrict__ nlow,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ perm,
                                                  const double* __restrict__ rD,
                                                  const double* __restrict__ cf,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ z, const KState* st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= ncell) return;
    const int c = cells[i];
    const double d = rD[c];
    double w = d * r[c];
    const int a = rowptr[c];
    const int e = a + nlow[c];
    for (int p = a; p < e; ++p) w = w - d * cf[perm[p]] * z[col[p]];
    z[c] = w;
}

__global__ __launch_bounds__(kBlk) void k_dic_bwd(int ncell, const int* __restrict__ cells,
                                                  const int* __restrict__ rowptr,
                                                  const int* __restrict__ nlow,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ perm, int base,
                                                  const double* __restrict__ rD,
                                                  const double* __restrict__ cb,
                                                  double* __restrict__ z, const KState* st) {
    if (st->done) return;
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= ncell) return;
    const int c = cells[i];
    const double d = rD[c];
    double w = z[c];
    const int b0 = rowptr[c] + nlow[c] + 1;   // after the diagonal: upper faces, face order
    for (int p = rowptr[c + 1] - 1; p >= b0; --p) w = w - d * cb[perm[p] - base] * z[col[p]];
    z[c] = w;
}

// rho = (r, z) after a DIC application; first: the set-up's rho, else also beta
template <int FIRST>
__global__ __launch_bounds__(kWG) void k_pcg_rz(int N, const double* __restrict__ r,
                                                const double* __restrict__ z, KState* st,
                                                RedWs red) {
    if (st->done) return;
    __shared__ RedShared shr;
    StreamRows t(N);
    double rv[4], zv[4];
    load4(r, t.i0, N, rv);
    load4(z, t.i0, N, zv);
    double leaf[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (t.i0 + k < N) leaf[0] = fma(rv[k], zv[k], leaf[0]);
    finish_tile_slots<1>(red, st, t.tile, leaf, shr, [&](const double (&v)[1]) {
        if (FIRST) {
            st->rho0 = v[0];
        } else {
            st->rho1 = st->rho0;
            st->rho0 = v[0];
            st->beta = st->rho0 / st->rho1;
        }
    });
}

// dependency levels of both sweeps from the face addressing (one pass each:
// faces are in owner order, so a face's lower cell is final when it is met
// going forward, its upper cell going backward), cells bucketed by level
int dic_levels(gis_context* h) {
    AmgData& A = *h->amg;
    const int N = h->N, M = h->M;
    std::vector<int> la, ua;
    GIS_TRY(download(h, la, h->lower_addr, (size_t)M));
    GIS_TRY(download(h, ua, h->upper_addr, (size_t)M));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    std::vector<int> lf(N, 0), lb(N, 0);
    for (int f = 0; f < M; ++f) lf[ua[f]] = std::max(lf[ua[f]], lf[la[f]] + 1);
    for (int f = M - 1; f >= 0; --f) lb[la[f]] = std::max(lb[la[f]], lb[ua[f]] + 1);
    auto group = [&](const std::vector<int>& lev, std::vector<int>& ptr, DevMem& cells) -> int {
        const int nl = N > 0 ? 1 + *std::max_element(lev.begin(), lev.end()) : 0;
        std::vector<int> all(N);
        for (int i = 0; i < N; ++i) all[i] = i;
        agglom::hvec<int> p, list;
        agglom::bucket(nl, lev, all, p, list);
        ptr.assign(p.begin(), p.end());
        return upload_vec(h, cells, list);
    };
    GIS_TRY(group(lf, A.dic_ptr_f, A.dic_cells_f));
    GIS_TRY(group(lb, A.dic_ptr_b, A.dic_cells_b));
    A.dic_levels = true;
    return kOK;
}

// kind 4: DIC (the factor from upper alone), 5: DILU (upper and lower)
int ilu_prepare(gis_context* h, int kind) {
    if (h->n_cpl > 0)
        return fail(h, kErrState, "DIC/DILU do not take interface entries (gis_set_interfaces)");
    if (h->batch.ready) return fail(h, kErrState, "DIC/DILU need a single-system pattern");
    if (!h->amg) {
        h->amg = new (std::nothrow) AmgData();
        if (!h->amg) return fail(h, kErrNoMem, "preconditioner data");
    }
    AmgData& A = *h->amg;
    if (!A.dic_levels) {
        int st = kOK;
        try {
            st = dic_levels(h);
        } catch (const std::bad_alloc&) {
            st = fail(h, kErrNoMem, "DIC/DILU set-up: host memory exhausted");
        } catch (const std::exception& e) {
            st = fail(h, kErrNoMem, std::string("DIC/DILU set-up: ") + e.what());
        }
        if (st != kOK) return st;
        A.dic_values = false;
    }
    if (!A.dic_values || A.dic_kind != kind) {
        const int N = h->N;
        GIS_HIP(h, A.dic_rd.ensure(sizeof(double) * (size_t)N));
        GIS_HIP(h, h->dinv.ensure(sizeof(double) * (size_t)N));
        double* F = h->dinv.as<double>();   // the factored diagonal before its reciprocal
        const int nl = (int)A.dic_ptr_f.size() - 1;
        for (int l = 0; l < nl; ++l) {
            const int n = A.dic_ptr_f[l + 1] - A.dic_ptr_f[l];
            hipLaunchKernelGGL(k_dic_factor, grid_of(n), dim3(kBlk), 0, h->stream, n,
                               A.dic_cells_f.as<int>() + A.dic_ptr_f[l], h->rowptr.as<int>(),
                               h->nbr_cnt.as<int>(), h->col.as<int>(), h->perm.as<int>(),
                               h->diag.as<double>(), h->upper.as<double>(),
                               kind == 5 ? h->lower.as<double>() : h->upper.as<double>(), F);
        }
        hipLaunchKernelGGL(k_recip, grid_of(N), dim3(kBlk), 0, h->stream, N, F,
                           A.dic_rd.as<double>());
        GIS_HIP(h, hipGetLastError());
        A.dic_values = true;
        A.dic_kind = kind;
    }
    return kOK;
}

int dic_prepare(gis_context* h) { return ilu_prepare(h, 4); }

// z = M^-1 r by the two sweeps, level by level: cf the forward sweep's
// coefficients (below the diagonal), cb the backward sweep's (above it)
int ilu_apply(gis_context* h, const double* r, double* z, const double* cf, const double* cb) {
    AmgData& A = *h->amg;
    const int nf = (int)A.dic_ptr_f.size() - 1, nb = (int)A.dic_ptr_b.size() - 1;
    for (int l = 0; l < nf; ++l) {
        const int n = A.dic_ptr_f[l + 1] - A.dic_ptr_f[l];
        hipLaunchKernelGGL(k_dic_fwd, grid_of(n), dim3(kBlk), 0, h->stream, n,
                           A.dic_cells_f.as<int>() + A.dic_ptr_f[l], h->rowptr.as<int>(),
                           h->nbr_cnt.as<int>(), h->col.as<int>(), h->perm.as<int>(),
                           A.dic_rd.as<double>(), cf, r, z, h->st);
    }
    for (int l = 0; l < nb; ++l) {
        const int n = A.dic_ptr_b[l + 1] - A.dic_ptr_b[l];
        hipLaunchKernelGGL(k_dic_bwd, grid_of(n), dim3(kBlk), 0, h->stream, n,
                           A.dic_cells_b.as<int>() + A.dic_ptr_b[l], h->rowptr.as<int>(),
                           h->nbr_cnt.as<int>(), h->col.as<int>(), h->perm.as<int>(),
                           h->M + h->N, A.dic_rd.as<double>(), cb, z, h->st);
    }
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int dic_apply(gis_context* h, const double* r, double* z) {
    return ilu_apply(h, r, z, h->upper.as<double>(), h->upper.as<double>());
}

int solve_pcg(gis_context* h, const SolveArgs& a, int pc, double* x, const double* b) {
    if (pc < 0 || pc > 4) return fail(h, kErrArg, "unknown preconditioner");
    if (pc == 2 || pc == 3) GIS_TRY(amg_prepare(h));
    if (pc == 4) GIS_TRY(dic_prepare(h));
    if (pc == 3 && !h->amg->values32_ok) GIS_TRY(amg_values32(h));
    GIS_TRY(prepare_solve(h, a, 5));
    const int N = h->N;
    double* r = h->vec.as<double>();
    double* pb0 = r + N;
    double* pb1 = r + 2 * (size_t)N;
    double* q = r + 3 * (size_t)N;
    double* z = r + 4 * (size_t)N;
    if (pc == 1) {
        GIS_HIP(h, h->dinv.ensure(sizeof(double) * (size_t)N));
        hipLaunchKernelGGL(k_recip, grid_of(N), dim3(kBlk), 0, h->stream, N,
                           h->diag.as<double>(), h->dinv.as<double>());
        GIS_HIP(h, hipGetLastError());
    }
    // PC 1: dinv = 1 / diag; PC 2: the finest level's sm (k_pcg2 writes the
    // V-cycle's first sweep sm .* r into z0)
    const double* dinv = pc == 1 ? h->dinv.as<double>()
                                 : (pc == 2 ? h->amg->lv[0].sm.as<double>() : nullptr);
    double* z0 = pc == 2 ? h->amg->lv[0].vecs.as<double>() : z;
    // PC 3: the finest level's FP32 rhs (r32) and first sweep (z32)
    float* z32 = pc == 3 ? h->amg->lv[0].vecs32.as<float>() : nullptr;
    float* r32 = pc == 3 ? z32 + 3 * (size_t)N : nullptr;
    const float* sm32 = pc == 3 ? h->amg->lv[0].sm32.as<float>() : nullptr;
    const double* zin = pc == 0 ? r : z;
    const dim3 g = tiles_grid(N);
    const Mat<double> m = mat64(h);
    GIS_SPMV_DISPATCH(h->sched.spmv64, P, {
        hipLaunchKernelGGL((k_setup<P>), g, dim3(kWG), 0, h->stream, m, N, b, x, nullptr, r,
                           nullptr, h->st, h->red, 0);
        GIS_HIP(h, hipGetLastError());
        if (pc == 1)
            hipLaunchKernelGGL(k_pcg_diag0, g, dim3(kWG), 0, h->stream, N, dinv, r, z, h->st,
                               h->red);
        if (pc == 2) {
            hipLaunchKernelGGL((k_vc_pre<double>), g, dim3(kWG), 0, h->stream, N, dinv, r, z0,
                               h->st);
            (void)vcycle<P, double>(h, 0, r, z, r);
        }
        if (pc == 3) {
            hipLaunchKernelGGL(k_vc_narrow_pre, g, dim3(kWG), 0, h->stream, N, sm32, r, r32, z32,
                               h->st);
            (void)vcycle<P, float>(h, 0, r32, z, r);
        }
        if (pc == 4) {
            GIS_TRY(dic_apply(h, r, z));
            hipLaunchKernelGGL(k_pcg_rz<1>, g, dim3(kWG), 0, h->stream, N, r, z, h->st, h->red);
        }
        GIS_HIP(h, hipGetLastError());
        GIS_COARSEN_DISPATCH(h->sched.coarsen64, CC, {
            const dim3 gc = super_grid(N, CC);
            GIS_TRY(run_blocks(h, iteration_budget(h), [&]() -> int {
                hipLaunchKernelGGL(k_pcg_dir, g, dim3(kWG), 0, h->stream, N, zin, pb0, pb1, x,
                                   h->st);
                hipLaunchKernelGGL((k_cg1s<P, CC>), gc, dim3(kWG), 0, h->stream, m, N, pb0, pb1,
                                   q, h->st, h->red);
                if (pc == 0) {
                    hipLaunchKernelGGL((k_pcg2<CC, 0>), gc, dim3(kWG), 0, h->stream, N, r, q,
                                       dinv, z, sm32, r32, z32, h->st, h->red);
                } else if (pc == 1) {
                    hipLaunchKernelGGL((k_pcg2<CC, 1>), gc, dim3(kWG), 0, h->stream, N, r, q,
                                       dinv, z, sm32, r32, z32, h->st, h->red);
                } else if (pc == 2) {
                    hipLaunchKernelGGL((k_pcg2<CC, 2>), gc, dim3(kWG), 0, h->stream, N, r, q,
                                       dinv, z0, sm32, r32, z32, h->st, h->red);
                    (void)vcycle<P, double>(h, 0, r, z, r);
                } else if (pc == 3) {
                    hipLaunchKernelGGL((k_pcg2<CC, 3>), gc, dim3(kWG), 0, h->stream, N, r, q,
                                       dinv, z, sm32, r32, z32, h->st, h->red);
                    (void)vcycle<P, float>(h, 0, r32, z, r);
                } else {
                    hipLaunchKernelGGL((k_pcg2<CC, 4>), gc, dim3(kWG), 0, h->stream, N, r, q,
                                       dinv, z, sm32, r32, z32, h->st, h->red);
                    GIS_TRY(dic_apply(h, r, z));
                    hipLaunchKernelGGL(k_pcg_rz<0>, g, dim3(kWG), 0, h->stream, N, r, z, h->st,
                                       h->red);
                }
                return kOK;
            }, pc == 2 || pc == 4));
        });
    });
    hipLaunchKernelGGL(k_cg_flush, g, dim3(kWG), 0, h->stream, N, x, pb0, pb1, h->st, 0);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int amg_levels(gis_context* h, int* n_levels) {
    GIS_TRY(amg_prepare(h));
    *n_levels = (int)h->amg->lv.size();
    return kOK;
}

int amg_level(gis_context* h, int l, int* n, int* m, int* agg, int* la, int* ua, double* diag,
              double* lower, double* upper) {
    GIS_TRY(amg_prepare(h));
    AmgData& A = *h->amg;
    if (l < 0 || l >= (int)A.lv.size()) return fail(h, kErrArg, "no such GAMG level");
    const AmgLevel& L = A.lv[l];
    gis_context* c = level_ctx(h, l);
    if (n) *n = L.n;
    if (m) *m = L.m;
    auto get = [&](void* dst, const DevMem& src, size_t bytes) -> int {
        if (dst && bytes)
            GIS_HIP(h, hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, h->stream));
        return kOK;
    };
    const size_t bm = (size_t)L.m, bn = (size_t)L.n;
    GIS_TRY(get(la, c->lower_addr, sizeof(int) * bm));
    GIS_TRY(get(ua, c->upper_addr, sizeof(int) * bm));
    GIS_TRY(get(diag, c->diag, sizeof(double) * bn));
    GIS_TRY(get(lower, c->lower, sizeof(double) * bm));
    GIS_TRY(get(upper, c->upper, sizeof(double) * bm));
    if (l + 1 < (int)A.lv.size()) GIS_TRY(get(agg, L.agg, sizeof(int) * bn));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return kOK;
}

}  // namespace gis
