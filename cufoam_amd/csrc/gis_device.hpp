// Device-side building blocks of the gis kernels for gfx950 (wave64).
//
// Arithmetic contract (identical in oracle/gis_oracle.c, so results can be
// compared bit for bit):
//   * compiled with -ffp-contract=off; fused multiply-adds only via fma()/fmaf();
//   * CSR row sums are sequential in storage order, products rounded first;
//   * every reduction has ONE order fixed by the vector length: leaves of
//     kLeaf = 4 consecutive elements accumulated from +0.0, then a pairwise
//     binary tree over the leaves.  A workgroup of kWG threads owns an aligned
//     tile of kWG*kLeaf rows (one subtree); partials of tiles are folded in
//     aligned groups of kGroup, level by level (SlotReduce below), so the
//     result never depends on the launch grid or the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gis {

constexpr int kWG = 256;              // threads per workgroup (4 wave64)
constexpr int kLeaf = 4;              // consecutive rows per reduction leaf
constexpr int kTile = kWG * kLeaf;    // rows per workgroup tile (one partial)
constexpr int kGroup = 64;            // partials folded per group (one wave)
constexpr int kMaxLevels = 6;         // kGroup^5 * kTile rows >> 2^31
constexpr int kMaxNV = 16;            // reductions carried by one kernel (8 RHS x 2)

// Cross-workgroup reduction workspace (device pointers, passed by value).
struct RedWs {
    double* part[kMaxLevels];         // level-L partials, count[L] * kMaxNV doubles
    double* slot;                     // published level-0 partials, count[0] * kMaxNV
                                      // doubles; every entry holds kSlotEmpty between launches
    double* slot1;                    // published group sums (level 1), count[1] * kMaxNV
    int count[kMaxLevels];            // number of partials at level L
    int nlevels;                      // levels until count == 1
};

// Solver scalars that live on the device between launches.
struct KState {
    double rho0, rho1, alpha, beta;   // Krylov scalars
    double norm, norm0;               // last and initial residual norm
    double error;                     // absolute stopping threshold
    double omega;                     // Richardson relaxation factor
    double int_bound;                 // BiCG breakdown bound
    double aux[8];                    // per-solver extras (norm factor, x_bar, ...)
    double* hist;                     // residual history (device), hist_cap entries
    long long max_steps;              // iteration cap (already made positive)
    int k;                            // completed iterations
    int done;                         // 1: stop, later launches are no-ops
    int status;                       // solve outcome code
    int hist_cap;
    int flag;                         // per-solver control word
    int nhist;                        // entries pushed by solvers with sparse histories
    double rel_tol;                   // > 0: error := max(error, rel_tol * norm0)
    long long min_steps;              // no stop on convergence before this many steps
    int slot_fault;                   // failure-path tests only: > 0 makes reduction
                                      // collectors see no slot written and give up
                                      // after that many polls
    int detail;                       // history: 0 residual norms, 1 elapsed seconds
    long long t0;                     // device wall clock when the loop started
    double tick;                      // seconds per device wall-clock tick
};

// OpenFOAM's loop rule (lduMatrix::solver): iterate while (k < maxIter and not
// converged) or k < minIter.  With min_steps = 0 it is the reference's rule.
__device__ __forceinline__ bool stop_after(const KState* st, bool converged) {
    return (converged || st->k >= st->max_steps) && st->k >= st->min_steps;
}

// ---------------------------------------------------------------------------
// cache-bypassing agent-scope accesses for the cross-workgroup partials.
// Agent-scope fences on gfx950 write back and invalidate the L2; the partials
// travel through atomics that bypass the non-coherent caches instead, and the
// fences below only order the memory operations of this wave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void st_agent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       static_cast<unsigned long long>(__double_as_longlong(v)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double ld_agent(const double* p) {
    return __longlong_as_double(static_cast<long long>(
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT)));
}

__device__ __forceinline__ void order_release() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));   // vmcnt(0)
}

__device__ __forceinline__ void order_acquire() {
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------------
// pairwise trees
// ---------------------------------------------------------------------------
// Butterfly over the 64 lanes: afterwards every lane holds the pairwise-tree
// sum of the wave's 64 values (addition is commutative, so partner lanes agree).
template <int NV>
__device__ __forceinline__ void wave_tree(double (&v)[NV]) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = v[i] + __shfl_xor(v[i], m, 64);
    }
}

// Pairwise tree over the kWG thread values; the result lands in thread 0.
template <int NV>
__device__ __forceinline__ void block_tree(double (&v)[NV], double (*sh)[kMaxNV]) {
    wave_tree<NV>(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[w][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = (sh[0][i] + sh[1][i]) + (sh[2][i] + sh[3][i]);
    }
}

// LDS a kernel hands to its reduction (SlotReduce, finish_tile_slots).
struct RedShared {
    double w[4][kMaxNV];
    double top[2 * kGroup];   // SlotReduce: the collecting workgroup's level-2 sums
    int last;
};

// ---------------------------------------------------------------------------
// Reduction without a per-workgroup round trip.
//
// Had each workgroup taken an arrival ticket (an atomic counter per group, the
// last arriver folding the group), it could not retire until the ticket's
// atomic came back, and the last arriver's fold would be a chain of such round
// trips at the very end of the kernel.  Here a workgroup only publishes its
// tiles' subtree sums, with plain agent-scope stores, into slots that hold the
// signalling-NaN pattern kSlotEmpty until written, and retires.  No arithmetic result is a
// signalling NaN (IEEE operations quiet them), so a slot that no longer holds
// the pattern holds its final value.
//   * The workgroup holding the last tile of an aligned group of kGroup tiles
//     collects that group's slots (one wave, one round trip), folds them with
//     wave_tree and publishes the group sum into a level-1 slot.
//   * The workgroup holding the last tile overall does the same for its own
//     group, then collects the level-1 slots and folds the levels above in
//     registers and LDS, and calls fin(totals).
// Every fold is a wave_tree over an aligned group of kGroup entries, zero-
// padded: the tree the oracle's reduction fixes, whatever the grid or the
// number of tiles per workgroup.  Collectors restore the
// pattern in every slot they read.  A collector waits only for workgroups that
// never wait themselves, and every wait is bounded: after kSlotSpins polls
// without the data it records kErrSync and stops the solve instead of hanging.
// ---------------------------------------------------------------------------
constexpr unsigned long long kSlotEmpty = 0x7FF5A5A5A5A5A5A5ull;   // a signalling NaN
constexpr int kSlotSpins = 1 << 20;
constexpr int kErrSync = 13;

__device__ __forceinline__ bool slot_empty(double v) {
    return static_cast<unsigned long long>(__double_as_longlong(v)) == kSlotEmpty;
}

__device__ __forceinline__ void slot_clear(double* p) {
    st_agent(p, __longlong_as_double(static_cast<long long>(kSlotEmpty)));
}

// The NV values of entry e (stride NV) of a slot array, once every lane of the
// wave sees its entry written; entries e >= n read as 0.  Each read slot is
// restored to the pattern.  False when the bounded wait ran out.
template <int NV>
__device__ __forceinline__ bool poll_entry(double* base, int e, int n, double (&x)[NV],
                                           int fault) {
    const int limit = fault > 0 ? fault : kSlotSpins;
    for (int spins = 0;; ++spins) {
        int miss = fault > 0 ? 1 : 0;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            x[i] = e < n ? ld_agent(base + (size_t)e * NV + i) : 0.0;
            miss |= slot_empty(x[i]) ? 1 : 0;
        }
        if (!__any(miss)) break;
        if (spins >= limit) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    if (e < n) {
#pragma unroll
        for (int i = 0; i < NV; ++i) slot_clear(base + (size_t)e * NV + i);
    }
    return true;
}

// A reduction whose bounded wait ran out: stop the solve(s) it serves.
__device__ __forceinline__ void slot_fail(KState* st, int nst) {
    for (int i = 0; i < nst; ++i) {
        st[i].status = kErrSync;
        st[i].done = 1;
    }
}

// The reduction of a kernel whose workgroups cover C consecutive tiles each:
//   begin(): the tiles' subtrees are formed and thread 0 publishes them; the
//            workgroup is then free to issue its output stores and retire;
//   end():   group and top collectors as described above; fin on thread 0 of
//            the top collector.
// With a single tile the totals go to fin at once.
template <int NV, int C>
struct SlotReduce {
    static_assert(NV <= kMaxNV, "at most kMaxNV totals");
    // level-2 sums fold in LDS when they fit RedShared::top
    static constexpr int kTopCap = 2 * kGroup;
    double flat[C * NV];
    int last_tile;

    __device__ void begin(const RedWs& red, int tile0, int nt, double (&v)[C][NV],
                          RedShared& sh) {
        static_assert(C * NV <= kMaxNV, "RedShared holds kMaxNV values per wave");
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int i = 0; i < NV; ++i) flat[c * NV + i] = v[c][i];
        block_tree<C * NV>(flat, sh.w);
        last_tile = nt > 0 ? tile0 + nt - 1 : -1;
        if (red.count[0] != 1 && threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (c < nt) {
#pragma unroll
                    for (int i = 0; i < NV; ++i)
                        st_agent(red.slot + (size_t)(tile0 + c) * NV + i, flat[c * NV + i]);
                }
            }
        }
    }

    // begin() for a kernel that formed its tiles' subtree sums itself (thread 0
    // holds them, e.g. pair_tile_sum): publish them as begin() would
    __device__ void begin_sums(const RedWs& red, int tile0, int nt, const double (&v)[C][NV]) {
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int i = 0; i < NV; ++i) flat[c * NV + i] = v[c][i];
        last_tile = nt > 0 ? tile0 + nt - 1 : -1;
        if (red.count[0] != 1 && threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (c < nt) {
#pragma unroll
                    for (int i = 0; i < NV; ++i)
                        st_agent(red.slot + (size_t)(tile0 + c) * NV + i, flat[c * NV + i]);
                }
            }
        }
    }

    // st[0 .. nst) are marked failed if a wait runs out
    template <class Fin>
    __device__ void end(const RedWs& red, KState* st, RedShared& sh, Fin&& fin, int nst = 1) {
        const int n0 = red.count[0];
        if (n0 == 1) {
            if (threadIdx.x == 0) {
                double t[NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) t[i] = flat[i];
                fin(t);
            }
            return;
        }
        const int t = last_tile;
        if (t < 0 || ((t + 1) % kGroup != 0 && t != n0 - 1)) return;
        const bool top = (t == n0 - 1);
        if (!top && threadIdx.x >= 64) return;
        const int lane = threadIdx.x & 63;
        const int g = t / kGroup;
        if (threadIdx.x < 64) {
            double v[NV];
            const bool ok = poll_entry<NV>(red.slot, g * kGroup + lane, n0, v, st->slot_fault);
            if (ok) wave_tree<NV>(v);
            if (!top) {
                if (ok && lane == 0) {
#pragma unroll
                    for (int i = 0; i < NV; ++i) st_agent(red.slot1 + (size_t)g * NV + i, v[i]);
                }
                return;
            }
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < NV; ++i) sh.w[0][i] = v[i];
                sh.last = ok ? 1 : 0;
            }
        }
        __syncthreads();
        top_collect(red, st, nst, sh, fin);
    }

    // the top collector, all threads: level-1 entries (its own group's sum
    // from sh.w[0], the others from the level-1 slots), then the levels above
    template <class Fin>
    __device__ void top_collect(const RedWs& red, KState* st, int nst, RedShared& sh,
                                Fin&& fin) {
        const int lane = threadIdx.x & 63;
        const int w = threadIdx.x >> 6;
        constexpr int kWaves = kWG / 64;
        const int n1 = red.count[1];
        const int n2 = red.count[2];
        double own[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) own[i] = sh.w[0][i];
        if (!sh.last) {
            if (threadIdx.x == 0) slot_fail(st, nst);
            return;
        }
        if (n1 == 1) {
            if (threadIdx.x == 0) fin(own);
            return;
        }
        bool ok = true;
        for (int h = w; h < n2; h += kWaves) {
            const int e = h * kGroup + lane;
            double x[NV];
            ok = poll_entry<NV>(red.slot1, e, n1 - 1, x, st->slot_fault) && ok;
            if (e == n1 - 1) {
#pragma unroll
                for (int i = 0; i < NV; ++i) x[i] = own[i];
            }
            wave_tree<NV>(x);
            if (n2 == 1) {
                if (ok && threadIdx.x == 0) fin(x);
            } else if (lane == 0) {
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    if (n2 <= kGroup && n2 * NV <= kTopCap)
                        sh.top[h * NV + i] = x[i];
                    else
                        st_agent(&red.part[2][(size_t)h * kMaxNV + i], x[i]);
                }
            }
        }
        if (n2 == 1) {
            if (!ok && threadIdx.x == 0) slot_fail(st, nst);
            return;
        }
        __syncthreads();
        if (!ok) sh.last = 0;   // only 0 is written here
        order_release();
        __syncthreads();
        if (!sh.last) {
            if (threadIdx.x == 0) slot_fail(st, nst);
            return;
        }
        if (n2 <= kGroup && n2 * NV <= kTopCap) {   // the next level is the total
            if (threadIdx.x < 64) {
                double x[NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) x[i] = lane < n2 ? sh.top[lane * NV + i] : 0.0;
                wave_tree<NV>(x);
                if (threadIdx.x == 0) fin(x);
            }
            return;
        }
        // level-2 sums that do not fit RedShared::top (many totals, or more than
        // kGroup^3 tiles): the levels above 2 through the global partials
        for (int L = 2; L < kMaxLevels; ++L) {
            const int n = red.count[L];
            if (n == 1) {
                if (threadIdx.x == 0) {
                    double x[NV];
#pragma unroll
                    for (int i = 0; i < NV; ++i) x[i] = ld_agent(&red.part[L][i]);
                    fin(x);
                }
                return;
            }
            order_acquire();
            const int groups = (n + kGroup - 1) / kGroup;
            for (int q = w; q < groups; q += kWaves) {
                double x[NV];
                const int src = q * kGroup + lane;
#pragma unroll
                for (int i = 0; i < NV; ++i)
                    x[i] = src < n ? ld_agent(&red.part[L][(size_t)src * kMaxNV + i]) : 0.0;
                wave_tree<NV>(x);
                if (lane == 0) {
#pragma unroll
                    for (int i = 0; i < NV; ++i)
                        st_agent(&red.part[L + 1][(size_t)q * kMaxNV + i], x[i]);
                }
            }
            order_release();
            __syncthreads();
        }
    }
};

// finish_tile over published slots (SlotReduce, one tile per workgroup): the
// same totals, passed to fin by one workgroup.  Every workgroup of the launch
// must call it.
template <int NV, class Fin>
__device__ __forceinline__ void finish_tile_slots(const RedWs& red, KState* st, int tile,
                                                  double (&v)[NV], RedShared& sh, Fin&& fin,
                                                  int nst = 1) {
    double vv[1][NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) vv[0][i] = v[i];
    SlotReduce<NV, 1> tr;
    tr.begin(red, tile, 1, vv, sh);
    tr.end(red, st, sh, fin, nst);
}


// ---------------------------------------------------------------------------
// CSR tile SpMV: y[row0 .. row0+nrows) = sum_j a_j * X(col_j), written to
// sh_y.  X(c) is evaluated on the fly (it may combine several vectors).
// Each row is summed sequentially from 0 in storage order.  Products are
// staged through LDS in chunks of at most CAP nonzeros so that val/col are
// read with unit stride across the workgroup.
// ---------------------------------------------------------------------------
template <typename T, int CAP>
struct SpmvShared {
    int rp[kTile + 1];
    T prod[CAP];
    int end;
};

template <typename T, int CAP, class XOp>
__device__ __forceinline__ void tile_spmv(const int* __restrict__ rowptr,
                                          const int* __restrict__ col,
                                          const T* __restrict__ val, int row0, int nrows,
                                          XOp&& xop, SpmvShared<T, CAP>& sh,
                                          T* __restrict__ sh_y) {
    const int tid = threadIdx.x;
    for (int i = tid; i <= nrows; i += kWG) sh.rp[i] = rowptr[row0 + i];
    __syncthreads();
    int r = 0;
    while (r < nrows) {
        if (tid == 0) {
            const int base = sh.rp[r];
            int e;
            if (sh.rp[r + 1] - base > CAP) {
                e = r;   // a single row longer than the staging buffer
            } else {
                int lo = r + 1, hi = nrows;   // invariant: rp[lo]-base <= CAP
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (sh.rp[mid] - base <= CAP) lo = mid; else hi = mid - 1;
                }
                e = lo;
            }
            sh.end = e;
        }
        __syncthreads();
        const int e = sh.end;
        if (e == r) {
            if (tid == 0) {
                T s = T(0);
                for (int j = sh.rp[r]; j < sh.rp[r + 1]; ++j) s = s + val[j] * xop(col[j]);
                sh_y[r] = s;
            }
            __syncthreads();
            r = r + 1;
            continue;
        }
        const int nz0 = sh.rp[r];
        const int cnt = sh.rp[e] - nz0;
#pragma unroll 4
        for (int j = tid; j < cnt; j += kWG) sh.prod[j] = val[nz0 + j] * xop(col[nz0 + j]);
        __syncthreads();
        for (int i = r + tid; i < e; i += kWG) {
            const int a = sh.rp[i] - nz0;
            const int b = sh.rp[i + 1] - nz0;
            T s = T(0);
            for (int j = a; j < b; ++j) s = s + sh.prod[j];
            sh_y[i] = s;
        }
        __syncthreads();
        r = e;
    }
}

// ---------------------------------------------------------------------------
// Sliced layout (SELL-64): rows are bundled 64 to a slice (one row per lane of
// a wave64); slice s stores its slots slot-major, slot j of lane i at
// soff[s] + 64*j + i, padding marked by column -1.  Slot j of a row is its
// j-th CSR entry, so the sequential row sum below is the CSR row sum bit for
// bit.  Every load of values and columns is unit-stride across the wave and
// the x gathers of neighbouring rows hit neighbouring addresses.
// A workgroup tile of kTile rows = 16 slices; wave w owns slices 4w..4w+3 and
// writes y to sh_y in row order (each wave touches only its own 256 rows).
// ---------------------------------------------------------------------------
constexpr int kSlice = 64;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// B = slots per batch (all of a batch's loads are issued before its gathers).
template <typename T, int B = 8, class XOp>
__device__ __forceinline__ T sell_row(const int* __restrict__ scol, const T* __restrict__ sval,
                                      int a, int len, int lane, XOp&& xop) {
    T acc = T(0);
    for (int j = 0; j < len; j += B) {
        int cc[B];
        T v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int idx = a + (j + u) * kSlice + lane;
            if (j + u < len) {
                cc[u] = __builtin_nontemporal_load(scol + idx);
                v[u] = __builtin_nontemporal_load(sval + idx);
            } else {
                cc[u] = -1;
                v[u] = T(0);
            }
        }
        T xv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) xv[u] = cc[u] >= 0 ? xop(cc[u]) : T(0);
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (cc[u] >= 0) acc = acc + v[u] * xv[u];
    }
    return acc;
}

// Two rows per lane from two slices at once: both slices' column and value
// loads are in flight together (twice the memory-level parallelism of
// sell_row per wave).  Each row is still summed sequentially in slot order.
template <typename T, class XOp>
__device__ __forceinline__ void sell_row2(const int* __restrict__ scol,
                                          const T* __restrict__ sval, int a0, int len0, int a1,
                                          int len1, int lane, XOp&& xop, T& y0, T& y1) {
    T acc0 = T(0), acc1 = T(0);
    const int len = len0 > len1 ? len0 : len1;
    for (int j = 0; j < len; j += 8) {
        int c0[8], c1[8];
        T v0[8], v1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int s = (j + u) * kSlice + lane;
            c0[u] = -1;
            v0[u] = T(0);
            c1[u] = -1;
            v1[u] = T(0);
            if (j + u < len0) {
                c0[u] = __builtin_nontemporal_load(scol + a0 + s);
                v0[u] = __builtin_nontemporal_load(sval + a0 + s);
            }
            if (j + u < len1) {
                c1[u] = __builtin_nontemporal_load(scol + a1 + s);
                v1[u] = __builtin_nontemporal_load(sval + a1 + s);
            }
        }
        T x0[8], x1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x0[u] = c0[u] >= 0 ? xop(c0[u]) : T(0);
            x1[u] = c1[u] >= 0 ? xop(c1[u]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (c0[u] >= 0) acc0 = acc0 + v0[u] * x0[u];
            if (c1[u] >= 0) acc1 = acc1 + v1[u] * x1[u];
        }
    }
    y0 = acc0;
    y1 = acc1;
}

// y[row0 .. row0+1024) of one tile into sh_y (tile row order).  Each wave sums
// 4 slices of 64 rows, one row per lane; KPAIR takes the 4 slices two at a
// time (sell_row2).
template <typename T, bool KPAIR, class XOp>
__device__ __forceinline__ void tile_sell_any(const int* __restrict__ soff,
                                              const int* __restrict__ scol,
                                              const T* __restrict__ sval, int N, int row0,
                                              XOp&& xop, T* __restrict__ sh_y) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
    if (KPAIR) {
#pragma unroll
        for (int c = 0; c < 4; c += 2) {
            const int s = s0 + c;
            int a0 = 0, len0 = 0, a1 = 0, len1 = 0;
            if (s < nslices) {
                a0 = soff[s];
                len0 = (soff[s + 1] - a0) / kSlice;
            }
            if (s + 1 < nslices) {
                a1 = soff[s + 1];
                len1 = (soff[s + 2] - a1) / kSlice;
            }
            T y0, y1;
            sell_row2<T>(scol, sval, a0, len0, a1, len1, lane, xop, y0, y1);
            sh_y[(w * 4 + c) * kSlice + lane] = y0;
            sh_y[(w * 4 + c + 1) * kSlice + lane] = y1;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int s = s0 + c;
            T y = T(0);
            if (s < nslices) {
                const int a = soff[s];
                const int len = (soff[s + 1] - a) / kSlice;
                y = sell_row<T>(scol, sval, a, len, lane, xop);
            }
            sh_y[(w * 4 + c) * kSlice + lane] = y;
        }
    }
    wave_sync_lds();
}

// One row per lane, compressed columns: slot j of the row sits in column
// row + sbase[a/64 + j] + sres[a + 64 j + lane] (residual -128: padding).  The
// base is one int per 64 slots, the residual one byte, so the index stream is
// about a quarter of sell_row's.  Same products, same order, same sum.
template <typename T, class XOp>
__device__ __forceinline__ T sell_row_c(const signed char* __restrict__ sres,
                                        const int* __restrict__ sbase,
                                        const T* __restrict__ sval, int a, int len, int lane,
                                        int row, XOp&& xop) {
    T acc = T(0);
    const int a64 = a / kSlice;
    for (int j = 0; j < len; j += 8) {
        int cc[8];
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = a + (j + u) * kSlice + lane;
            cc[u] = -1;
            v[u] = T(0);
            if (j + u < len) {
                const int r = sres[idx];
                const int b = sbase[a64 + j + u];
                cc[u] = (r == -128) ? -1 : row + b + r;
                v[u] = __builtin_nontemporal_load(sval + idx);
            }
        }
        T xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = cc[u] >= 0 ? xop(cc[u]) : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (cc[u] >= 0) acc = acc + v[u] * xv[u];
    }
    return acc;
}

// tile_sell over compressed columns; slices whose residuals do not fit in a
// byte (sflag 0) read the int32 columns.
template <typename T, class XOp>
__device__ __forceinline__ void tile_sell_c(const int* __restrict__ soff,
                                            const int* __restrict__ scol,
                                            const signed char* __restrict__ sres,
                                            const int* __restrict__ sbase,
                                            const unsigned char* __restrict__ sflag,
                                            const T* __restrict__ sval, int N, int row0,
                                            XOp&& xop, T* __restrict__ sh_y) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int s = s0 + c;
        T y = T(0);
        if (s < nslices) {
            const int a = soff[s];
            const int len = (soff[s + 1] - a) / kSlice;
            if (sflag[s])
                y = sell_row_c<T>(sres, sbase, sval, a, len, lane, s * kSlice + lane, xop);
            else
                y = sell_row<T>(scol, sval, a, len, lane, xop);
        }
        sh_y[(w * 4 + c) * kSlice + lane] = y;
    }
    wave_sync_lds();
}

// Slices with a column-offset table (SellD): the slice's rows share a sorted
// list of U column offsets (doffs[ooff[s] ..]); slot k of a row holds column
// row + doffs[ooff[s] + k] when bit k of the row's mask is set, and is a
// padding slot otherwise.  A row's entries ascend in column, so the set bits
// visit them in storage order: same products, same order, same sum as
// sell_row, with no column index per slot.  Values: slot-major from 64 * ooff[s].
//
// The loop body has no per-slot branches: slot indices are clamped into the
// slice (min(k, U-1)), every lane gathers column row + offset clamped into
// [0, last] whatever its mask bit, and a clear bit leaves the sum untouched
// through a select.  The gather addresses therefore depend only on the
// (scalar) offset table, not on the mask: values, masks and gathers of a batch
// are all in flight together, one memory round trip per batch.
template <typename T, class XOp>
__device__ __forceinline__ T sell_row_d(const int* __restrict__ offs, unsigned mask,
                                        const T* __restrict__ dval, int a, int len, int lane,
                                        int row, int last, XOp&& xop) {
    T acc = T(0);
    for (int j = 0; j < len; j += 8) {
        T v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = min(j + u, len - 1);
            v[u] = __builtin_nontemporal_load(dval + a + k * kSlice + lane);
        }
        int o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) o[u] = offs[min(j + u, len - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xop(min(max(row + o[u], 0), last));
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool on = (mask >> (j + u)) & 1u;
            const T t = acc + v[u] * xv[u];
            acc = on ? t : acc;
        }
    }
    return acc;
}

// Two offset-table slices per lane at a time (their loads in flight together).
template <typename T, class XOp>
__device__ __forceinline__ void sell_row2_d(const int* __restrict__ offs0, unsigned m0, int a0,
                                            int len0, int row0,
                                            const int* __restrict__ offs1, unsigned m1, int a1,
                                            int len1, int row1, int last,
                                            const T* __restrict__ dval, int lane, XOp&& xop,
                                            T& y0, T& y1) {
    T acc0 = T(0), acc1 = T(0);
    const int len = len0 > len1 ? len0 : len1;
    for (int j = 0; j < len; j += 8) {
        T v0[8], v1[8], x0[8], x1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k0 = min(j + u, len0 - 1), k1 = min(j + u, len1 - 1);
            v0[u] = __builtin_nontemporal_load(dval + a0 + k0 * kSlice + lane);
            v1[u] = __builtin_nontemporal_load(dval + a1 + k1 * kSlice + lane);
        }
        int o0[8], o1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            o0[u] = offs0[min(j + u, len0 - 1)];
            o1[u] = offs1[min(j + u, len1 - 1)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x0[u] = xop(min(max(row0 + o0[u], 0), last));
            x1[u] = xop(min(max(row1 + o1[u], 0), last));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool on0 = (m0 >> (j + u)) & 1u, on1 = (m1 >> (j + u)) & 1u;
            const T t0 = acc0 + v0[u] * x0[u];
            const T t1 = acc1 + v1[u] * x1[u];
            acc0 = on0 ? t0 : acc0;
            acc1 = on1 ? t1 : acc1;
        }
    }
    y0 = acc0;
    y1 = acc1;
}

// tile over SellD slices two at a time; a slice without a table (ooff step
// 0: its rows do not ascend, or their offsets do not fit 32 slots) reads the
// int32 slot columns instead.
// Row mask of a table slice: one byte per row where the table has at most 8
// entries (3 bytes a row less to stream), the 32-bit mask otherwise.  U is
// uniform over the wave, so only one of the two arrays is read.
__device__ __forceinline__ unsigned row_mask(const unsigned* __restrict__ dmask,
                                             const unsigned char* __restrict__ dmask8, int U,
                                             int row, int N) {
    unsigned m = 0u;
    if (U <= 8) {
        if (row < N) m = dmask8[row];
    } else {
        if (row < N) m = dmask[row];
    }
    return m;
}

template <typename T, class XOp>
__device__ __forceinline__ void tile_sell_d(const int* __restrict__ soff,
                                            const int* __restrict__ scol,
                                            const T* __restrict__ sval,
                                            const int* __restrict__ ooff,
                                            const int* __restrict__ doffs,
                                            const unsigned* __restrict__ dmask,
                                            const unsigned char* __restrict__ dmask8,
                                            const T* __restrict__ dval, int N, int row0,
                                            XOp&& xop, T* __restrict__ sh_y) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
#pragma unroll
    for (int c = 0; c < 4; c += 2) {
        const int s = s0 + c;
        int u0 = 0, u1 = 0, o0 = 0, o1 = 0;
        if (s < nslices) {
            o0 = ooff[s];
            u0 = ooff[s + 1] - o0;
        }
        if (s + 1 < nslices) {
            o1 = ooff[s + 1];
            u1 = ooff[s + 2] - o1;
        }
        T y0 = T(0), y1 = T(0);
        const int r0 = s * kSlice + lane, r1 = (s + 1) * kSlice + lane;
        // rows past N have an empty mask; their (unused) gathers stay inside [0, N)
        if (u0 > 0 && u1 > 0) {
            const unsigned m0 = row_mask(dmask, dmask8, u0, r0, N);
            const unsigned m1 = row_mask(dmask, dmask8, u1, r1, N);
            sell_row2_d<T>(doffs + o0, m0, o0 * kSlice, u0, r0, doffs + o1, m1, o1 * kSlice,
                           u1, r1, N - 1, dval, lane, xop, y0, y1);
        } else {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int sp = s + p;
                if (sp >= nslices) continue;
                const int up = p ? u1 : u0, op = p ? o1 : o0, rp = p ? r1 : r0;
                T y;
                if (up > 0) {
                    const unsigned mp = row_mask(dmask, dmask8, up, rp, N);
                    y = sell_row_d<T>(doffs + op, mp, dval, op * kSlice, up, lane, rp, N - 1,
                                      xop);
                } else {
                    const int a = soff[sp];
                    y = sell_row<T>(scol, sval, a, (soff[sp + 1] - a) / kSlice, lane, xop);
                }
                if (p) y1 = y; else y0 = y;
            }
        }
        sh_y[(w * 4 + c) * kSlice + lane] = y0;
        sh_y[(w * 4 + c + 1) * kSlice + lane] = y1;
    }
    wave_sync_lds();
}

// Symmetric storage of a SellD layout (a bitwise symmetric matrix, every slice
// tabled): only the slots of table entries with offset >= 0 are stored (dsv).
// Each table entry has one 16-byte record e = {offset o, base a, base b, 0}
// (one scalar load per slot: three separate scalar loads serialised the loop).
// Slot k of a row reads dsv[base + (r & 63)], r = row + min(o, 0): its own slot
// for o >= 0 (r = row, base = a), the +d slot of row r = row - d for o = -d,
// whose slice is the one of row 64 s - d (base a) or the next (base b).  Record
// and slice test are uniform per slot, so the loop keeps the branch-free form
// of sell_row2_d: same values, order and sum.  The loads are cached: each
// stored upper slot is read a second time as a lower slot, by the same wave
// (d < 64), a nearby one, or the workgroup 64 tiles on.
template <typename T>
__device__ __forceinline__ T sym_load(const T* __restrict__ dsv, const int4& e, int srow, int row) {
    const int so = min(e.x, 0);
    const int r = row + so;
    const int s1 = (srow + so) >> 6;   // floor
    const int base = (r >> 6) == s1 ? e.y : e.z;
    return dsv[base + (r & 63)];
}

template <typename T, class XOp>
__device__ __forceinline__ void sell_row2_ds(const int4* __restrict__ tab0, unsigned m0, int len0,
                                             int srow0, const int4* __restrict__ tab1,
                                             unsigned m1, int len1, int srow1, int lane, int last,
                                             const T* __restrict__ dsv, XOp&& xop, T& y0, T& y1) {
    T acc0 = T(0), acc1 = T(0);
    const int row0 = srow0 + lane, row1 = srow1 + lane;
    const int len = len0 > len1 ? len0 : len1;
    for (int j = 0; j < len; j += 8) {
        T v0[8], v1[8], x0[8], x1[8];
        int4 e0[8], e1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            e0[u] = tab0[min(j + u, len0 - 1)];
            e1[u] = tab1[min(j + u, len1 - 1)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v0[u] = sym_load(dsv, e0[u], srow0, row0);
            v1[u] = sym_load(dsv, e1[u], srow1, row1);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x0[u] = xop(min(max(row0 + e0[u].x, 0), last));
            x1[u] = xop(min(max(row1 + e1[u].x, 0), last));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool on0 = (m0 >> (j + u)) & 1u, on1 = (m1 >> (j + u)) & 1u;
            const T t0 = acc0 + v0[u] * x0[u];
            const T t1 = acc1 + v1[u] * x1[u];
            acc0 = on0 ? t0 : acc0;
            acc1 = on1 ? t1 : acc1;
        }
    }
    y0 = acc0;
    y1 = acc1;
}

template <typename T, class XOp>
__device__ __forceinline__ T sell_row_ds(const int4* __restrict__ tab, unsigned mask, int len,
                                         int srow, int lane, int last,
                                         const T* __restrict__ dsv, XOp&& xop) {
    T acc = T(0);
    const int row = srow + lane;
    for (int j = 0; j < len; j += 8) {
        T v[8], xv[8];
        int4 e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) e[u] = tab[min(j + u, len - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = sym_load(dsv, e[u], srow, row);
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xop(min(max(row + e[u].x, 0), last));
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool on = (mask >> (j + u)) & 1u;
            const T t = acc + v[u] * xv[u];
            acc = on ? t : acc;
        }
    }
    return acc;
}

// tile over the slices two at a time in the symmetric storage (every slice
// has a table; slices past the end are skipped)
template <typename T, class XOp>
__device__ __forceinline__ void tile_sell_ds(const int* __restrict__ ooff,
                                             const unsigned* __restrict__ dmask,
                                             const unsigned char* __restrict__ dmask8,
                                             const int4* __restrict__ symx,
                                             const T* __restrict__ dsv, int N, int row0,
                                             XOp&& xop, T* __restrict__ sh_y) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nslices = (N + kSlice - 1) / kSlice;
    const int s0 = row0 / kSlice + w * 4;
#pragma unroll
    for (int c = 0; c < 4; c += 2) {
        const int s = s0 + c;
        T y0 = T(0), y1 = T(0);
        if (s + 1 < nslices) {
            const int o0 = ooff[s], o1 = ooff[s + 1];
            const int u0 = o1 - o0, u1 = ooff[s + 2] - o1;
            const unsigned m0 = row_mask(dmask, dmask8, u0, s * kSlice + lane, N);
            const unsigned m1 = row_mask(dmask, dmask8, u1, (s + 1) * kSlice + lane, N);
            sell_row2_ds<T>(symx + o0, m0, u0, s * kSlice, symx + o1, m1, u1, (s + 1) * kSlice,
                            lane, N - 1, dsv, xop, y0, y1);
        } else if (s < nslices) {
            const int o0 = ooff[s];
            const int u0 = ooff[s + 1] - o0;
            const unsigned m0 = row_mask(dmask, dmask8, u0, s * kSlice + lane, N);
            y0 = sell_row_ds<T>(symx + o0, m0, u0, s * kSlice, lane, N - 1, dsv, xop);
        }
        sh_y[(w * 4 + c) * kSlice + lane] = y0;
        sh_y[(w * 4 + c + 1) * kSlice + lane] = y1;
    }
    wave_sync_lds();
}

template <typename T, class XOp>
__device__ __forceinline__ void tile_sell(const int* __restrict__ soff,
                                          const int* __restrict__ scol,
                                          const T* __restrict__ sval, int N, int row0,
                                          XOp&& xop, T* __restrict__ sh_y) {
    tile_sell_any<T, false>(soff, scol, sval, N, row0, xop, sh_y);
}

template <typename T, class XOp>
__device__ __forceinline__ void tile_sell2(const int* __restrict__ soff,
                                           const int* __restrict__ scol,
                                           const T* __restrict__ sval, int N, int row0,
                                           XOp&& xop, T* __restrict__ sh_y) {
    tile_sell_any<T, true>(soff, scol, sval, N, row0, xop, sh_y);
}

// Unstaged variant: each thread sums its own kLeaf consecutive rows straight
// from global memory (same order, same result; fewer LDS passes, strided loads).
template <typename T, class XOp>
__device__ __forceinline__ void tile_spmv_direct(const int* __restrict__ rowptr,
                                                 const int* __restrict__ col,
                                                 const T* __restrict__ val, int row0,
                                                 int nrows, XOp&& xop, T* __restrict__ sh_y) {
    const int t0 = threadIdx.x * kLeaf;
#pragma unroll
    for (int k = 0; k < kLeaf; ++k) {
        const int i = t0 + k;
        if (i < nrows) {
            const int a = rowptr[row0 + i];
            const int b = rowptr[row0 + i + 1];
            T s = T(0);
            for (int j = a; j < b; ++j) s = s + val[j] * xop(col[j]);
            sh_y[i] = s;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// kLeaf consecutive elements owned by one thread
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load4(const double* __restrict__ p, int i0, int n, double (&o)[4]) {
    if (i0 + 3 < n) {
        const double2 a = *reinterpret_cast<const double2*>(p + i0);
        const double2 b = *reinterpret_cast<const double2*>(p + i0 + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (i0 + k < n) ? p[i0 + k] : 0.0;
    }
}

__device__ __forceinline__ void store4(double* __restrict__ p, int i0, int n, const double (&o)[4]) {
    if (i0 + 3 < n) {
        *reinterpret_cast<double2*>(p + i0) = make_double2(o[0], o[1]);
        *reinterpret_cast<double2*>(p + i0 + 2) = make_double2(o[2], o[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i0 + k < n) p[i0 + k] = o[k];
    }
}

// Streaming (non-temporal) stores for the output of a kernel that gathers x:
// the fresh output lines then do not evict the gathered lines from the L2.
// Whole rows go out as 16-byte vector stores, as store4 does.
typedef double gis_d2 __attribute__((ext_vector_type(2)));
typedef float gis_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void store4nt(double* __restrict__ p, int i0, int n,
                                         const double (&o)[4]) {
    if (i0 + 3 < n) {
        gis_d2 a, b;
        a.x = o[0]; a.y = o[1];
        b.x = o[2]; b.y = o[3];
        __builtin_nontemporal_store(a, reinterpret_cast<gis_d2*>(p + i0));
        __builtin_nontemporal_store(b, reinterpret_cast<gis_d2*>(p + i0 + 2));
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i0 + k < n) p[i0 + k] = o[k];
    }
}

__device__ __forceinline__ void store4nt(float* __restrict__ p, int i0, int n,
                                         const float (&o)[4]) {
    if (i0 + 3 < n) {
        gis_f4 a;
        a.x = o[0]; a.y = o[1]; a.z = o[2]; a.w = o[3];
        __builtin_nontemporal_store(a, reinterpret_cast<gis_f4*>(p + i0));
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i0 + k < n) p[i0 + k] = o[k];
    }
}

// A tile's kTile FP64 results from LDS (tile row order) to y with streaming
// stores that write whole cache lines per instruction: thread t stores rows
// 2t, 2t+1 and 512+2t, 513+2t of the tile as 16-byte vectors, consecutive
// threads on consecutive rows.  (Per-thread rows 4t..4t+3 as two 16-byte
// stores would leave every 128-byte line half-written by each instruction,
// and streaming stores reach memory as those partial lines.)  The caller has
// made sh_y complete with a workgroup barrier.
__device__ __forceinline__ void store_tile_nt(double* __restrict__ y,
                                              const double* __restrict__ sh_y, int row0, int n) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lr = 2 * (int)threadIdx.x + h * (kTile / 2);
        const int g = row0 + lr;
        if (g + 1 < n) {
            gis_d2 v;
            v.x = sh_y[lr];
            v.y = sh_y[lr + 1];
            __builtin_nontemporal_store(v, reinterpret_cast<gis_d2*>(y + g));
        } else if (g < n) {
            y[g] = sh_y[lr];
        }
    }
}

// The same through the caches, for the FP64 CG's q (read by k_cg2 next): with
// the symmetric storage, 380.1 -> 378.0 us per FP64 CG iteration at 256^3
__device__ __forceinline__ void store_tile_cached(double* __restrict__ y,
                                                  const double* __restrict__ sh_y, int row0,
                                                  int n) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lr = 2 * (int)threadIdx.x + h * (kTile / 2);
        const int g = row0 + lr;
        if (g + 1 < n) {
            gis_d2 v;
            v.x = sh_y[lr];
            v.y = sh_y[lr + 1];
            *reinterpret_cast<gis_d2*>(y + g) = v;
        } else if (g < n) {
            y[g] = sh_y[lr];
        }
    }
}

// Coalesced rows of a tile for the FP64 CG's streaming kernels: thread t owns
// rows row0 + 2t, +1 (o[0], o[1]) and row0 + 512 + 2t, +1 (o[2], o[3]), so each
// 16-byte access of a wave covers 1 KB of consecutive rows.  Loads and stores
// are non-temporal: these vectors are touched once per kernel, and keeping them
// out of the caches leaves the caches to the SpMV's gathers (measured at 256^3:
// 450 -> 418 us per FP64 CG iteration).
__device__ __forceinline__ void load_pairs(const double* __restrict__ p, int row0, int n,
                                           double (&o)[4]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = row0 + 2 * (int)threadIdx.x + h * (kTile / 2);
        if (g + 1 < n) {
            const gis_d2 a = __builtin_nontemporal_load(reinterpret_cast<const gis_d2*>(p + g));
            o[2 * h] = a.x;
            o[2 * h + 1] = a.y;
        } else {
            o[2 * h] = g < n ? p[g] : 0.0;
            o[2 * h + 1] = 0.0;
        }
    }
}

__device__ __forceinline__ void store_pairs(double* __restrict__ p, int row0, int n,
                                            const double (&o)[4]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = row0 + 2 * (int)threadIdx.x + h * (kTile / 2);
        if (g + 1 < n) {
            gis_d2 v;
            v.x = o[2 * h];
            v.y = o[2 * h + 1];
            __builtin_nontemporal_store(v, reinterpret_cast<gis_d2*>(p + g));
        } else if (g < n) {
            p[g] = o[2 * h];
        }
    }
}

// The same rows stored through the caches, for a vector the next kernel reads
// at once (the FP64 CG's r: k_cg2 74.7 -> 67.4 us at 256^3 against store_pairs).
__device__ __forceinline__ void store_pairs_cached(double* __restrict__ p, int row0, int n,
                                                   const double (&o)[4]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = row0 + 2 * (int)threadIdx.x + h * (kTile / 2);
        if (g + 1 < n) {
            gis_d2 v;
            v.x = o[2 * h];
            v.y = o[2 * h + 1];
            *reinterpret_cast<gis_d2*>(p + g) = v;
        } else if (g < n) {
            p[g] = o[2 * h];
        }
    }
}

// The tile's subtree sum of squares sum v_i^2 for values held in the pair
// layout of load_pairs (thread t: rows 2t, 2t+1 and 512+2t, 513+2t), in the
// canonical order: a leaf (rows 4j..4j+3, accumulated from +0.0, rows >= n
// skipped) is started by the even lane of a pair and finished by the odd one;
// each wave's 32 leaves of a half tile are folded pairwise across the odd
// lanes; the halves' wave sums are combined ((w0 + w1) + (w2 + w3)) and the two
// halves added.  These are the bits block_tree gives for per-thread leaves of
// four rows.  The result is in thread 0; every thread of the workgroup calls it.
__device__ __forceinline__ double pair_tile_sumsq(const double (&v)[4], int row0, int n,
                                                  RedShared& sh) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    double half[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = row0 + 2 * (int)threadIdx.x + h * (kTile / 2);
        double a = 0.0;
        if (g < n) a = fma(v[2 * h], v[2 * h], a);
        if (g + 1 < n) a = fma(v[2 * h + 1], v[2 * h + 1], a);
        const double first = __shfl_xor(a, 1, 64);   // the even lane's two rows
        double leaf = first;
        if (g < n) leaf = fma(v[2 * h], v[2 * h], leaf);
        if (g + 1 < n) leaf = fma(v[2 * h + 1], v[2 * h + 1], leaf);
#pragma unroll
        for (int m = 2; m < 64; m <<= 1) leaf = leaf + __shfl_xor(leaf, m, 64);
        half[h] = leaf;   // in the odd lanes: the sum of this wave's 32 leaves
    }
    if (lane == 1) {
        sh.w[w][0] = half[0];
        sh.w[w][1] = half[1];
    }
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const double s0 = (sh.w[0][0] + sh.w[1][0]) + (sh.w[2][0] + sh.w[3][0]);
        const double s1 = (sh.w[0][1] + sh.w[1][1]) + (sh.w[2][1] + sh.w[3][1]);
        s = s0 + s1;
    }
    __syncthreads();   // sh.w is reused by the reduction that follows
    return s;
}

__device__ __forceinline__ void load4(const float* __restrict__ p, int i0, int n, float (&o)[4]) {
    if (i0 + 3 < n) {
        const float4 a = *reinterpret_cast<const float4*>(p + i0);
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (i0 + k < n) ? p[i0 + k] : 0.0f;
    }
}

// Non-temporal load4 (a stream read once, as k_mx1's correction xc)
__device__ __forceinline__ void load4nt(const float* __restrict__ p, int i0, int n, float (&o)[4]) {
    if (i0 + 3 < n) {
        const gis_f4 a = __builtin_nontemporal_load(reinterpret_cast<const gis_f4*>(p + i0));
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (i0 + k < n) ? p[i0 + k] : 0.0f;
    }
}

__device__ __forceinline__ void store4(float* __restrict__ p, int i0, int n, const float (&o)[4]) {
    if (i0 + 3 < n) {
        *reinterpret_cast<float4*>(p + i0) = make_float4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i0 + k < n) p[i0 + k] = o[k];
    }
}

}  // namespace gis
