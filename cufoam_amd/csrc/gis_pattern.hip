// Device LDU -> CSR conversion (replaces the host loop of gis/ldu2csr.h:7-42).
//
// Row c of the CSR holds, in this order (gis/ldu2csr.h:19-39):
//   lower faces with neighbour c, ascending face index  (col = owner,     val = lower[f])
//   the diagonal                                        (col = c,         val = diag[c])
//   upper faces owned by c, ascending face index        (col = neighbour, val = upper[f])
// so row_offsets[c] = c + losortStart[c] + ownerStart[c].
//
// The pattern is built once per mesh (integer work only, bit-exact and
// deterministic: atomics only count or claim slots that a per-row sort then
// orders); a per-nonzero code `perm` remembers where each CSR value comes
// from, so every later coefficient update is one gather:
//   perm[p] in [0, M)      -> lower[f]
//   perm[p] in [M, M+N)    -> diag[c]
//   perm[p] in [M+N, 2M+N) -> upper[f]
// A^T of an LDU matrix has the same pattern with lower and upper swapped.
#include <climits>

#include "gis_internal.hpp"

namespace gis {

namespace {

constexpr int kBlk = 256;

__global__ __launch_bounds__(kBlk) void k_count(int N, int M, const int* __restrict__ la,
                                                const int* __restrict__ ua, int* own_cnt,
                                                int* nbr_cnt, int* first_own, int* err) {
    const int f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= M) return;
    const int l = la[f];
    const int u = ua[f];
    bool bad = l < 0 || l >= N || u < 0 || u >= N;
    const int lp = f > 0 ? la[f - 1] : -1;
    if (f > 0 && lp > l) bad = true;   // owners must be sorted (upper-triangular order)
    if (bad) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(&own_cnt[l], 1);
    atomicAdd(&nbr_cnt[u], 1);
    if (lp != l) first_own[l] = f;
}

// coupling entries per row; row == col or an index out of range is an error
__global__ __launch_bounds__(kBlk) void k_count_cpl(int N, int C, const int* __restrict__ cr,
                                                    const int* __restrict__ cc, int* cpl_cnt,
                                                    int* err) {
    const int j = blockIdx.x * kBlk + threadIdx.x;
    if (j >= C) return;
    const int r = cr[j];
    const int c = cc[j];
    if (r < 0 || r >= N || c < 0 || c >= N || r == c) {
        atomicOr(err, 2);
        return;
    }
    atomicAdd(&cpl_cnt[r], 1);
}

__global__ __launch_bounds__(kBlk) void k_rowlen(int N, const int* __restrict__ own_cnt,
                                                 const int* __restrict__ nbr_cnt,
                                                 const int* __restrict__ cpl_cnt, int* rowlen) {
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c < N) rowlen[c] = 1 + own_cnt[c] + nbr_cnt[c] + cpl_cnt[c];
    else if (c == N) rowlen[c] = 0;
}

// exclusive scan of 1024-element blocks (256 threads x 4)
__global__ __launch_bounds__(kBlk) void k_scan_block(const int* __restrict__ in, int* out,
                                                     int* block_sums, int n) {
    __shared__ int sh[kBlk / 64 + 1];
    const int base = blockIdx.x * (kBlk * 4) + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < kBlk / 64; ++i) {
            const int t = sh[i];
            sh[i] = acc;
            acc += t;
        }
        sh[kBlk / 64] = acc;
    }
    __syncthreads();
    int run = x - s + sh[w];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = sh[kBlk / 64];
}

__global__ __launch_bounds__(kBlk) void k_scan_add(int* out, const int* __restrict__ offs, int n) {
    const int i = blockIdx.x * (kBlk * 4) + threadIdx.x * 4;
    const int o = offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + k < n) out[i + k] += o;
}

__global__ __launch_bounds__(kBlk) void k_fill_diag(int N, int M, const int* __restrict__ rowptr,
                                                    const int* __restrict__ nbr_cnt, int* col,
                                                    int* perm) {
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= N) return;
    const int pos = rowptr[c] + nbr_cnt[c];
    col[pos] = c;
    perm[pos] = M + c;
}

__global__ __launch_bounds__(kBlk) void k_fill_faces(int N, int M, const int* __restrict__ la,
                                                     const int* __restrict__ ua,
                                                     const int* __restrict__ rowptr,
                                                     const int* __restrict__ nbr_cnt,
                                                     const int* __restrict__ first_own, int* fill,
                                                     int* col, int* perm) {
    const int f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= M) return;
    const int o = la[f];
    const int u = ua[f];
    const int up = rowptr[o] + nbr_cnt[o] + 1 + (f - first_own[o]);
    col[up] = u;
    perm[up] = M + N + f;
    const int slot = atomicAdd(&fill[u], 1);
    perm[rowptr[u] + slot] = f;   // lower part, ordered by k_sort_lower
}

__global__ __launch_bounds__(kBlk) void k_sort_lower(int N, const int* __restrict__ la,
                                                     const int* __restrict__ rowptr,
                                                     const int* __restrict__ nbr_cnt, int* col,
                                                     int* perm) {
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= N) return;
    const int a = rowptr[c];
    const int n = nbr_cnt[c];
    for (int i = 1; i < n; ++i) {   // insertion sort by face index (rows are short)
        const int v = perm[a + i];
        int j = i - 1;
        while (j >= 0 && perm[a + j] > v) {
            perm[a + j + 1] = perm[a + j];
            --j;
        }
        perm[a + j + 1] = v;
    }
    for (int i = 0; i < n; ++i) col[a + i] = la[perm[a + i]];
}

// SellD tables, one wave per 64-row slice.  A slice gets a table when every
// row's columns strictly ascend, no row holds more than 32 entries, and the
// union of the rows' column offsets (col - row) has at most 32 entries and
// pads the slice by at most a quarter.  ucount[s] = table size (0: none).
__global__ __launch_bounds__(kBlk) void k_selld_union(int N, int S,
                                                      const int* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const int* __restrict__ soff, int* ucount,
                                                      int* tmp_offs, unsigned* dmask,
                                                      unsigned char* dmask8, int* counts) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (kBlk / 64) + (threadIdx.x >> 6);
    if (s >= S) return;   // whole waves
    const int row = s * 64 + lane;
    const bool valid = row < N;
    const int a = valid ? rowptr[row] : 0;
    const int len = valid ? rowptr[row + 1] - a : 0;
    bool ok = len <= 32;
    for (int j = 1; j < len && ok; ++j)
        if (col[a + j] <= col[a + j - 1]) ok = false;
    int bad = ok ? 0 : 1, total = len;
    for (int o = 32; o > 0; o >>= 1) {
        bad += __shfl_xor(bad, o, 64);
        total += __shfl_xor(total, o, 64);
    }
    bool fit = bad == 0;
    unsigned mask = 0u;
    int U = 0, p = 0;
    while (fit) {
        const int head = p < len ? col[a + p] - row : INT_MAX;
        int m = head;
        for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
        if (m == INT_MAX) break;
        if (U == 32) {
            fit = false;
            break;
        }
        if (head == m) {
            mask |= 1u << U;
            ++p;
        }
        if (lane == 0) tmp_offs[(size_t)s * 32 + U] = m;
        ++U;
    }
    if (fit && U * 64 * 4 > 5 * total) fit = false;
    if (valid) {
        dmask[row] = fit ? mask : 0u;
        dmask8[row] = (unsigned char)(fit && U <= 8 ? mask : 0u);
    }
    // counts: [0] slices without a table, [1] tabled slices with one-byte masks,
    // [2] sliced-row slots of the slices without a table
    if (lane == 0) {
        ucount[s] = fit ? U : 0;
        if (!fit) {
            atomicAdd(&counts[0], 1);
            atomicAdd(&counts[2], soff[s + 1] - soff[s]);
        } else if (U <= 8) {
            atomicAdd(&counts[1], 1);
        }
    }
}

// offsets compacted to doffs[ooff[s] ..]; slot (table entry k, lane) -> value
// code of the row's next entry when bit k is set, -1 for a padding slot
__global__ __launch_bounds__(kBlk) void k_selld_fill(int N, int S,
                                                     const int* __restrict__ rowptr,
                                                     const int* __restrict__ perm,
                                                     const int* __restrict__ ooff,
                                                     const int* __restrict__ tmp_offs,
                                                     const unsigned* __restrict__ dmask,
                                                     int* doffs, int* dperm) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (kBlk / 64) + (threadIdx.x >> 6);
    if (s >= S) return;
    const int o = ooff[s];
    const int U = ooff[s + 1] - o;
    if (U == 0) return;
    if (lane < U) doffs[o + lane] = tmp_offs[(size_t)s * 32 + lane];
    const int row = s * 64 + lane;
    const unsigned m = row < N ? dmask[row] : 0u;
    const int a = row < N ? rowptr[row] : 0;
    int j = 0;
    for (int k = 0; k < U; ++k) {
        int code = -1;
        if ((m >> k) & 1u) code = perm[a + j++];
        dperm[((size_t)o + k) * 64 + lane] = code;
    }
}

// coupling entries at the end of their row, then ordered by entry index
__global__ __launch_bounds__(kBlk) void k_fill_cpl(int N, int M, int C,
                                                   const int* __restrict__ cr,
                                                   const int* __restrict__ rowptr,
                                                   const int* __restrict__ cpl_cnt, int* cfill,
                                                   int* perm) {
    const int j = blockIdx.x * kBlk + threadIdx.x;
    if (j >= C) return;
    const int r = cr[j];
    const int slot = atomicAdd(&cfill[r], 1);
    perm[rowptr[r + 1] - cpl_cnt[r] + slot] = 2 * M + N + j;
}

__global__ __launch_bounds__(kBlk) void k_sort_cpl(int N, int M, const int* __restrict__ cc,
                                                   const int* __restrict__ rowptr,
                                                   const int* __restrict__ cpl_cnt, int* col,
                                                   int* perm) {
    const int r = blockIdx.x * kBlk + threadIdx.x;
    if (r >= N) return;
    const int n = cpl_cnt[r];
    const int a = rowptr[r + 1] - n;
    for (int i = 1; i < n; ++i) {
        const int v = perm[a + i];
        int j = i - 1;
        while (j >= 0 && perm[a + j] > v) {
            perm[a + j + 1] = perm[a + j];
            --j;
        }
        perm[a + j + 1] = v;
    }
    for (int i = 0; i < n; ++i) col[a + i] = cc[perm[a + i] - 2 * M - N];
}

// value of one CSR / slot position from its code (-1: padding slot); codes
// [0, M) lower, [M, M+N) diag, [M+N, 2M+N) upper, then coupling entries, read
// through cpl_map when it is given (the mirrored entry, for A^T)
template <typename T>
__global__ __launch_bounds__(kBlk) void k_gather(long long nnz, int N, int M,
                                                 const int* __restrict__ perm,
                                                 const double* __restrict__ diag,
                                                 const double* __restrict__ lower,
                                                 const double* __restrict__ upper,
                                                 const double* __restrict__ cpl,
                                                 const int* __restrict__ cpl_map, T* val) {
    const long long p = (long long)blockIdx.x * kBlk + threadIdx.x;
    if (p >= nnz) return;
    const int code = perm[p];
    double v;
    if (code < 0) v = 0.0;
    else if (code < M) v = lower[code];
    else if (code < M + N) v = diag[code - M];
    else if (code < 2 * M + N) v = upper[code - M - N];
    else {
        const int j = code - 2 * M - N;
        v = cpl[cpl_map ? cpl_map[j] : j];
    }
    val[p] = static_cast<T>(v);
}

// sliced layout: slot count of each 64-row slice = 64 * (longest row)
// Slots of each 64-row slice (64 x its longest row) and their 64-bit total: a
// skewed pattern (a few long rows spread over many slices) can pad past 2^31
// slots although nnz < 2^31; the host rejects it before the int32 scan.
__global__ __launch_bounds__(kBlk) void k_slice_len(int N, int S, const int* __restrict__ rowptr,
                                                    int* slots, unsigned long long* total) {
    const int s = blockIdx.x * kBlk + threadIdx.x;
    if (s > S) return;
    if (s == S) {
        slots[s] = 0;
        return;
    }
    int len = 0;
    const int r1 = min(N, (s + 1) * kSlice);
    for (int r = s * kSlice; r < r1; ++r) len = max(len, rowptr[r + 1] - rowptr[r]);
    const long long n = (long long)len * kSlice;
    slots[s] = (int)min(n, (long long)INT_MAX);
    atomicAdd(total, (unsigned long long)n);
}

__global__ __launch_bounds__(kBlk) void k_sell_fill(int N, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const int* __restrict__ perm,
                                                    const int* __restrict__ soff, int* scol,
                                                    int* sperm) {
    const int r = blockIdx.x * kBlk + threadIdx.x;
    if (r >= N) return;
    const int s = r / kSlice;
    const int lane = r % kSlice;
    const int a = rowptr[r];
    const int len = rowptr[r + 1] - a;
    const int base = soff[s] + lane;
    for (int j = 0; j < len; ++j) {
        scol[base + j * kSlice] = col[a + j];
        sperm[base + j * kSlice] = perm[a + j];
    }
}

// Compressed slot columns, one wave per slice.  For each slot row u the
// column offsets d = col - row of the valid lanes get a base B = midpoint of
// their range, and each lane stores d - B in one byte (-128: padding).  The
// slice is flagged usable only if every slot row's range spans at most 254.
__global__ __launch_bounds__(kSlice) void k_sell_compress(int S, const int* __restrict__ soff,
                                                          const int* __restrict__ scol,
                                                          signed char* __restrict__ sres,
                                                          int* __restrict__ sbase,
                                                          unsigned char* __restrict__ sflag) {
    const int s = blockIdx.x;
    if (s >= S) return;
    const int lane = threadIdx.x;
    const int a = soff[s];
    const int len = (soff[s + 1] - a) / kSlice;
    const long long row = (long long)s * kSlice + lane;
    int ok = 1;
    for (int u = 0; u < len; ++u) {
        const int idx = a + u * kSlice + lane;
        const int c = scol[idx];
        const bool valid = c >= 0;
        const long long d = valid ? (long long)c - row : 0;
        long long lo = valid ? d : LLONG_MAX;
        long long hi = valid ? d : LLONG_MIN;
#pragma unroll
        for (int m = 1; m < kSlice; m <<= 1) {
            lo = min(lo, (long long)__shfl_xor(lo, m, kSlice));
            hi = max(hi, (long long)__shfl_xor(hi, m, kSlice));
        }
        long long base = 0;
        if (lo <= hi) {
            if (hi - lo > 254) ok = 0;
            base = lo + (hi - lo) / 2;
        }
        const bool fits = ok && base >= INT_MIN && base <= INT_MAX;
        if (!fits) ok = 0;
        const long long r = valid ? d - base : -128;
        sres[idx] = (signed char)(ok ? r : -128);
        if (lane == 0) sbase[a / kSlice + u] = ok ? (int)base : 0;
    }
    if (lane == 0) sflag[s] = (unsigned char)ok;
}

// ---------------------------------------------------------------------------
// Symmetric storage of a SellD layout whose slices all have a table.  A
// bitwise symmetric matrix keeps the slots of its table entries with column
// offset >= 0 only: dsv, slot-major per slice from vb[s] = 64 x (entries with
// offset >= 0 of the earlier slices).  The entry of offset -d of row i is the
// entry of offset +d of row i - d, in the slice (i - d) / 64 at lane
// (i - d) % 64.  Per table entry e of slice s (offset o):
//   o >= 0: syma[e] = symb[e] = vb[s] + 64 x (e's rank among the entries >= 0);
//   o = -d: syma[e] = base of the +d entry in slice S1 = floor((64 s - d) / 64),
//           symb[e] = the same in slice S1 + 1 (the rows i - d of the slice lie
//           in these two); 0 where that slice does not exist or has no +d entry
//           (no row there holds the entry, so no set mask bit reads it).
// The values a row sums, their order and so the result are those of the full
// layout; k_sym_check confirms the symmetry on the values before use.
// ---------------------------------------------------------------------------

// slots of slice s that the symmetric storage keeps (64 per entry with offset >= 0)
__global__ __launch_bounds__(kBlk) void k_sym_count(int S, const int* __restrict__ ooff,
                                                    const int* __restrict__ doffs, int* cnt) {
    const int s = blockIdx.x * kBlk + threadIdx.x;
    if (s > S) return;
    int c = 0;
    if (s < S)
        for (int e = ooff[s]; e < ooff[s + 1]; ++e) c += doffs[e] >= 0 ? kSlice : 0;
    cnt[s] = c;
}

__device__ __forceinline__ int sym_base_of(int S, const int* __restrict__ ooff,
                                           const int* __restrict__ doffs,
                                           const int* __restrict__ vb, int sp, int d) {
    if (sp < 0 || sp >= S) return 0;
    int rank = 0;
    for (int e = ooff[sp]; e < ooff[sp + 1]; ++e) {
        const int o = doffs[e];
        if (o == d) return vb[sp] + rank * kSlice;
        if (o >= 0) ++rank;
    }
    return 0;
}

__global__ __launch_bounds__(kBlk) void k_sym_fill(int S, const int* __restrict__ ooff,
                                                   const int* __restrict__ doffs,
                                                   const int* __restrict__ vb, int* syma,
                                                   int* symb) {
    const int s = blockIdx.x * kBlk + threadIdx.x;
    if (s >= S) return;
    int rank = 0;
    for (int e = ooff[s]; e < ooff[s + 1]; ++e) {
        const int o = doffs[e];
        if (o >= 0) {
            syma[e] = symb[e] = vb[s] + rank * kSlice;
            ++rank;
        } else {
            const int s1 = (s * kSlice + o) >> 6;   // floor: o < 0 may take it below 0
            syma[e] = sym_base_of(S, ooff, doffs, vb, s1, -o);
            symb[e] = sym_base_of(S, ooff, doffs, vb, s1 + 1, -o);
        }
    }
}

// per table entry the record the SpMV reads: {offset, syma, symb, 0}
__global__ __launch_bounds__(kBlk) void k_sym_pack(int n_tab, const int* __restrict__ doffs,
                                                   const int* __restrict__ syma,
                                                   const int* __restrict__ symb, int4* symx) {
    const int e = blockIdx.x * kBlk + threadIdx.x;
    if (e < n_tab) symx[e] = make_int4(doffs[e], syma[e], symb[e], 0);
}

// dsv: the kept slots of the full slot values (one thread per full slot)
template <typename T>
__global__ __launch_bounds__(kBlk) void k_sym_compact(long long n_dslots,
                                                      const int* __restrict__ doffs,
                                                      const int* __restrict__ syma,
                                                      const T* __restrict__ dval, T* dsv) {
    const long long i = (long long)blockIdx.x * kBlk + threadIdx.x;
    if (i >= n_dslots) return;
    const int e = (int)(i / kSlice);
    if (doffs[e] >= 0) dsv[syma[e] + (int)(i % kSlice)] = dval[i];
}

// bad = 1 unless every set slot of an entry with offset < 0 holds the same bits
// as the slot the symmetric storage reads for it.  One wave per slice.
__global__ __launch_bounds__(kBlk) void k_sym_check(int N, int S, const int* __restrict__ ooff,
                                                    const int* __restrict__ doffs,
                                                    const unsigned* __restrict__ dmask,
                                                    const int* __restrict__ syma,
                                                    const int* __restrict__ symb,
                                                    const double* __restrict__ dval,
                                                    const double* __restrict__ dsv, int* bad) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (kBlk / 64) + (threadIdx.x >> 6);
    if (s >= S) return;   // whole waves
    const int row = s * kSlice + lane;
    const unsigned m = row < N ? dmask[row] : 0u;
    int mismatch = 0;
    for (int e = ooff[s], k = 0; e < ooff[s + 1]; ++e, ++k) {
        const int o = doffs[e];
        if (o >= 0) break;   // offsets ascend
        if (!((m >> k) & 1u)) continue;
        const int r = row + o;
        const int s1 = (s * kSlice + o) >> 6;
        const int base = (r >> 6) == s1 ? syma[e] : symb[e];
        if (__double_as_longlong(dval[(long long)e * kSlice + lane]) !=
            __double_as_longlong(dsv[base + (r & 63)]))
            mismatch = 1;
    }
    const unsigned long long any = __ballot(mismatch);
    if (lane == 0 && any) atomicOr(bad, 1);
}

int scan_exclusive(gis_context* h, const int* in, int* out, int n, int* tmp, size_t tmp_ints) {
    const int nb = (n + kBlk * 4 - 1) / (kBlk * 4);
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kBlk), 0, h->stream, in, out, nullptr, n);
        GIS_HIP(h, hipGetLastError());
        return kOK;
    }
    if ((size_t)2 * nb > tmp_ints) return fail(h, kErrNoMem, "scan scratch too small");
    int* sums = tmp;
    int* sums_scanned = tmp + nb;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kBlk), 0, h->stream, in, out, sums, n);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(scan_exclusive(h, sums, sums_scanned, nb, tmp + 2 * nb, tmp_ints - 2 * nb));
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlk), 0, h->stream, out, sums_scanned, n);
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

inline dim3 grid_for(long long n) { return dim3((unsigned)((n + kBlk - 1) / kBlk)); }

}  // namespace

int build_pattern(gis_context* h, int N, int M, const int* d_la, const int* d_ua) {
    const int C = h->n_cpl;
    const long long nnz = (long long)N + 2LL * M + C;
    if (N < 1 || M < 0 || C < 0 || nnz >= (1LL << 31))
        return fail(h, kErrShape,
                    "n_cells must be >= 1 and n_cells + 2*n_faces + n_couples < 2^31");
    h->has_pattern = false;
    h->has_values = h->has_val32 = h->has_valT = h->has_val32T = false;
    if (h->amg) h->amg->clear();
    GIS_HIP(h, h->rowptr.ensure(sizeof(int) * ((size_t)N + 1)));
    GIS_HIP(h, h->col.ensure(sizeof(int) * (size_t)nnz));
    GIS_HIP(h, h->perm.ensure(sizeof(int) * (size_t)nnz));
    GIS_HIP(h, h->nbr_cnt.ensure(sizeof(int) * (size_t)N));
    // scratch: own_cnt[N] first_own[N] fill[N] rowlen[N+1] err[1] cpl_cnt[N] cfill[N]
    // + scan temporaries
    const size_t scan_tmp = 4 * ((size_t)N / (kBlk * 4) + 64);
    const size_t scratch_ints = 6 * (size_t)N + 2 + scan_tmp;
    GIS_HIP(h, h->scratch.ensure(sizeof(int) * scratch_ints));
    int* own_cnt = h->scratch.as<int>();
    int* first_own = own_cnt + N;
    int* fill = first_own + N;
    int* rowlen = fill + N;
    int* err = rowlen + N + 1;
    int* cpl_cnt = err + 1;
    int* cfill = cpl_cnt + N;
    int* tmp = cfill + N;
    int* nbr_cnt = h->nbr_cnt.as<int>();
    GIS_HIP(h, hipMemsetAsync(own_cnt, 0, sizeof(int) * (3 * (size_t)N + 2), h->stream));
    GIS_HIP(h, hipMemsetAsync(err, 0, sizeof(int) * (2 * (size_t)N + 1), h->stream));
    GIS_HIP(h, hipMemsetAsync(nbr_cnt, 0, sizeof(int) * (size_t)N, h->stream));
    if (M > 0) {
        hipLaunchKernelGGL(k_count, grid_for(M), dim3(kBlk), 0, h->stream, N, M, d_la, d_ua,
                           own_cnt, nbr_cnt, first_own, err);
        GIS_HIP(h, hipGetLastError());
    }
    if (C > 0) {
        hipLaunchKernelGGL(k_count_cpl, grid_for(C), dim3(kBlk), 0, h->stream, N, C,
                           h->cpl_row.as<int>(), h->cpl_col.as<int>(), cpl_cnt, err);
        GIS_HIP(h, hipGetLastError());
    }
    hipLaunchKernelGGL(k_rowlen, grid_for((long long)N + 1), dim3(kBlk), 0, h->stream, N, own_cnt,
                       nbr_cnt, cpl_cnt, rowlen);
    GIS_HIP(h, hipGetLastError());
    GIS_TRY(scan_exclusive(h, rowlen, h->rowptr.as<int>(), N + 1, tmp, scan_tmp));
    int herr = 0;
    GIS_HIP(h, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (herr & 1)
        return fail(h, kErrArg,
                    "face addressing invalid: need 0 <= lowerAddr, upperAddr < n_cells and "
                    "lowerAddr sorted ascending");
    if (herr & 2)
        return fail(h, kErrArg, "interface entries need 0 <= row, col < n_cells and row != col");
    hipLaunchKernelGGL(k_fill_diag, grid_for(N), dim3(kBlk), 0, h->stream, N, M,
                       h->rowptr.as<int>(), nbr_cnt, h->col.as<int>(), h->perm.as<int>());
    GIS_HIP(h, hipGetLastError());
    if (M > 0) {
        hipLaunchKernelGGL(k_fill_faces, grid_for(M), dim3(kBlk), 0, h->stream, N, M, d_la, d_ua,
                           h->rowptr.as<int>(), nbr_cnt, first_own, fill, h->col.as<int>(),
                           h->perm.as<int>());
        GIS_HIP(h, hipGetLastError());
        hipLaunchKernelGGL(k_sort_lower, grid_for(N), dim3(kBlk), 0, h->stream, N, d_la,
                           h->rowptr.as<int>(), nbr_cnt, h->col.as<int>(), h->perm.as<int>());
        GIS_HIP(h, hipGetLastError());
    }
    if (C > 0) {
        hipLaunchKernelGGL(k_fill_cpl, grid_for(C), dim3(kBlk), 0, h->stream, N, M, C,
                           h->cpl_row.as<int>(), h->rowptr.as<int>(), cpl_cnt, cfill,
                           h->perm.as<int>());
        GIS_HIP(h, hipGetLastError());
        hipLaunchKernelGGL(k_sort_cpl, grid_for(N), dim3(kBlk), 0, h->stream, N, M,
                           h->cpl_col.as<int>(), h->rowptr.as<int>(), cpl_cnt, h->col.as<int>(),
                           h->perm.as<int>());
        GIS_HIP(h, hipGetLastError());
    }
    // sliced copy of the pattern (64-row slices, slot-major)
    const int S = (N + kSlice - 1) / kSlice;
    GIS_HIP(h, h->soff.ensure(sizeof(int) * ((size_t)S + 1)));
    int* slots = rowlen;   // reuse: S + 1 <= N + 1
    // own_cnt and first_own (adjacent) are free once the rows are filled: the
    // first 8 bytes hold the total
    unsigned long long* slot_total = reinterpret_cast<unsigned long long*>(own_cnt);
    GIS_HIP(h, hipMemsetAsync(slot_total, 0, sizeof(unsigned long long), h->stream));
    hipLaunchKernelGGL(k_slice_len, grid_for((long long)S + 1), dim3(kBlk), 0, h->stream, N, S,
                       h->rowptr.as<int>(), slots, slot_total);
    GIS_HIP(h, hipGetLastError());
    unsigned long long total_slots = 0;
    GIS_HIP(h, hipMemcpyAsync(&total_slots, slot_total, sizeof(total_slots),
                              hipMemcpyDeviceToHost, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (total_slots >= (1ULL << 31))
        return fail(h, kErrShape,
                    "sliced layout needs 2^31 or more slots (64 x the longest row of each "
                    "64-row slice): pattern too skewed");
    GIS_TRY(scan_exclusive(h, slots, h->soff.as<int>(), S + 1, tmp, scan_tmp));
    int n_slots = 0;
    GIS_HIP(h, hipMemcpyAsync(&n_slots, h->soff.as<int>() + S, sizeof(int),
                              hipMemcpyDeviceToHost, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    GIS_HIP(h, h->scol.ensure(sizeof(int) * (size_t)n_slots));
    GIS_HIP(h, h->sperm.ensure(sizeof(int) * (size_t)n_slots));
    GIS_HIP(h, hipMemsetAsync(h->scol.p, 0xff, sizeof(int) * (size_t)n_slots, h->stream));
    GIS_HIP(h, hipMemsetAsync(h->sperm.p, 0xff, sizeof(int) * (size_t)n_slots, h->stream));
    hipLaunchKernelGGL(k_sell_fill, grid_for(N), dim3(kBlk), 0, h->stream, N,
                       h->rowptr.as<int>(), h->col.as<int>(), h->perm.as<int>(),
                       h->soff.as<int>(), h->scol.as<int>(), h->sperm.as<int>());
    GIS_HIP(h, hipGetLastError());
    GIS_HIP(h, h->sres.ensure((size_t)n_slots));
    GIS_HIP(h, h->sbase.ensure(sizeof(int) * ((size_t)n_slots / kSlice + 1)));
    GIS_HIP(h, h->sflag.ensure((size_t)S + 1));
    hipLaunchKernelGGL(k_sell_compress, dim3((unsigned)S), dim3(kSlice), 0, h->stream, S,
                       h->soff.as<int>(), h->scol.as<int>(), h->sres.as<signed char>(),
                       h->sbase.as<int>(), h->sflag.as<unsigned char>());
    GIS_HIP(h, hipGetLastError());
    // column-offset tables (SellD)
    {
        DevMem ucount, tmp_offs;
        // ucount[0..S]: table sizes (ucount[S] = 0 for the scan); ucount[S+1..S+3]:
        // the layout counts of k_selld_union
        GIS_HIP(h, ucount.ensure(sizeof(int) * ((size_t)S + 4)));
        GIS_HIP(h, tmp_offs.ensure(sizeof(int) * 32 * (size_t)S));
        GIS_HIP(h, h->ooff.ensure(sizeof(int) * ((size_t)S + 1)));
        GIS_HIP(h, h->dmask.ensure(sizeof(unsigned) * (size_t)N));
        GIS_HIP(h, h->dmask8.ensure((size_t)N));
        GIS_HIP(h, hipMemsetAsync(ucount.p, 0, sizeof(int) * ((size_t)S + 4), h->stream));
        const dim3 gs((unsigned)((S + kBlk / 64 - 1) / (kBlk / 64)));
        hipLaunchKernelGGL(k_selld_union, gs, dim3(kBlk), 0, h->stream, N, S,
                           h->rowptr.as<int>(), h->col.as<int>(), h->soff.as<int>(),
                           ucount.as<int>(), tmp_offs.as<int>(), h->dmask.as<unsigned>(),
                           h->dmask8.as<unsigned char>(), ucount.as<int>() + S + 1);
        GIS_HIP(h, hipGetLastError());
        GIS_TRY(scan_exclusive(h, ucount.as<int>(), h->ooff.as<int>(), S + 1, tmp, scan_tmp));
        int n_tab = 0, counts[3] = {0, 0, 0};
        GIS_HIP(h, hipMemcpyAsync(&n_tab, h->ooff.as<int>() + S, sizeof(int),
                                  hipMemcpyDeviceToHost, h->stream));
        GIS_HIP(h, hipMemcpyAsync(counts, ucount.as<int>() + S + 1, sizeof(counts),
                                  hipMemcpyDeviceToHost, h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));
        h->lay_slices = S;
        h->lay_table_slices = S - counts[0];
        h->lay_table_entries = n_tab;
        h->lay_byte_mask_slices = counts[1];
        h->lay_untabled_slots = counts[2];
        h->n_dslots = 64LL * n_tab;
        GIS_HIP(h, h->doffs.ensure(sizeof(int) * std::max(n_tab, 1)));
        GIS_HIP(h, h->dperm.ensure(sizeof(int) * (size_t)std::max<long long>(h->n_dslots, 1)));
        hipLaunchKernelGGL(k_selld_fill, gs, dim3(kBlk), 0, h->stream, N, S, h->rowptr.as<int>(),
                           h->perm.as<int>(), h->ooff.as<int>(), tmp_offs.as<int>(),
                           h->dmask.as<unsigned>(), h->doffs.as<int>(), h->dperm.as<int>());
        GIS_HIP(h, hipGetLastError());
        GIS_HIP(h, hipStreamSynchronize(h->stream));   // the temporaries are freed next
    }
    // symmetric storage: only where every slice has a table
    h->sym_layout = false;
    h->sym_values = false;
    h->n_sym_slots = 0;
    if (S > 0 && h->lay_table_slices == S && h->n_dslots < (1LL << 31)) {
        DevMem cnt;
        GIS_HIP(h, cnt.ensure(sizeof(int) * ((size_t)S + 1)));
        GIS_HIP(h, h->syma.ensure(sizeof(int) * (size_t)h->lay_table_entries));
        GIS_HIP(h, h->symb.ensure(sizeof(int) * (size_t)h->lay_table_entries));
        hipLaunchKernelGGL(k_sym_count, grid_for((long long)S + 1), dim3(kBlk), 0, h->stream, S,
                           h->ooff.as<int>(), h->doffs.as<int>(), cnt.as<int>());
        GIS_HIP(h, hipGetLastError());
        DevMem vb;
        GIS_HIP(h, vb.ensure(sizeof(int) * ((size_t)S + 1)));
        GIS_TRY(scan_exclusive(h, cnt.as<int>(), vb.as<int>(), S + 1, tmp, scan_tmp));
        hipLaunchKernelGGL(k_sym_fill, grid_for(S), dim3(kBlk), 0, h->stream, S,
                           h->ooff.as<int>(), h->doffs.as<int>(), vb.as<int>(), h->syma.as<int>(),
                           h->symb.as<int>());
        GIS_HIP(h, hipGetLastError());
        const int n_tab = (int)h->lay_table_entries;
        GIS_HIP(h, h->symx.ensure(sizeof(int4) * (size_t)n_tab));
        hipLaunchKernelGGL(k_sym_pack, grid_for(n_tab), dim3(kBlk), 0, h->stream, n_tab,
                           h->doffs.as<int>(), h->syma.as<int>(), h->symb.as<int>(),
                           h->symx.as<int4>());
        GIS_HIP(h, hipGetLastError());
        int n_sym = 0;
        GIS_HIP(h, hipMemcpyAsync(&n_sym, vb.as<int>() + S, sizeof(int), hipMemcpyDeviceToHost,
                                  h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));   // cnt and vb are freed next
        // every base + 63 must lie in the kept slots, the fallback base 0 included
        if (n_sym >= kSlice) {
            h->n_sym_slots = n_sym;
            h->sym_layout = true;
        }
    }
    h->n_slots = n_slots;
    h->N = N;
    h->M = M;
    h->nnz = nnz;
    h->has_pattern = true;
    return kOK;
}

// CSR values and slot values of one matrix: A, or A^T (lower and upper swapped,
// each coupling entry taking the value of its mirrored entry)
template <typename T>
static int gather_both(gis_context* h, bool transpose, DevMem& csr, DevMem& sell,
                       DevMem& dsell) {
    const double* lower = (transpose ? h->upper : h->lower).as<double>();
    const double* upper = (transpose ? h->lower : h->upper).as<double>();
    const double* cpl = h->n_cpl ? h->cpl_val.as<double>() : nullptr;
    const int* map = (h->n_cpl && transpose) ? h->cpl_partner.as<int>() : nullptr;
    GIS_HIP(h, csr.ensure(sizeof(T) * (size_t)h->nnz));
    GIS_HIP(h, sell.ensure(sizeof(T) * (size_t)h->n_slots));
    hipLaunchKernelGGL(k_gather<T>, grid_for(h->nnz), dim3(kBlk), 0, h->stream, h->nnz, h->N,
                       h->M, h->perm.as<int>(), h->diag.as<double>(), lower, upper, cpl, map,
                       csr.as<T>());
    GIS_HIP(h, hipGetLastError());
    hipLaunchKernelGGL(k_gather<T>, grid_for(h->n_slots), dim3(kBlk), 0, h->stream, h->n_slots,
                       h->N, h->M, h->sperm.as<int>(), h->diag.as<double>(), lower, upper, cpl,
                       map, sell.as<T>());
    GIS_HIP(h, hipGetLastError());
    GIS_HIP(h, dsell.ensure(sizeof(T) * (size_t)std::max<long long>(h->n_dslots, 1)));
    if (h->n_dslots > 0)
        hipLaunchKernelGGL(k_gather<T>, grid_for(h->n_dslots), dim3(kBlk), 0, h->stream,
                           h->n_dslots, h->N, h->M, h->dperm.as<int>(), h->diag.as<double>(),
                           lower, upper, cpl, map, dsell.as<T>());
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

// Symmetric storage of T values: dsv from the full slot values
template <typename T>
static int sym_compact(gis_context* h, const DevMem& dval, DevMem& dsv) {
    GIS_HIP(h, dsv.ensure(sizeof(T) * (size_t)h->n_sym_slots));
    hipLaunchKernelGGL(k_sym_compact<T>, grid_for(h->n_dslots), dim3(kBlk), 0, h->stream,
                       h->n_dslots, h->doffs.as<int>(), h->syma.as<int>(), dval.as<T>(),
                       dsv.as<T>());
    GIS_HIP(h, hipGetLastError());
    return kOK;
}

int gather_values(gis_context* h) {
    if (!h->has_pattern) return fail(h, kErrState, "pattern not set");
    GIS_TRY(gather_both<double>(h, false, h->val, h->sval, h->dval));
    h->sym_values = false;
    if (h->sym_layout) {
        GIS_TRY(sym_compact<double>(h, h->dval, h->dsv));
        DevMem flag;
        GIS_HIP(h, flag.ensure(sizeof(int)));
        GIS_HIP(h, hipMemsetAsync(flag.p, 0, sizeof(int), h->stream));
        const int S = (int)h->lay_slices;
        hipLaunchKernelGGL(k_sym_check, dim3((unsigned)((S + kBlk / 64 - 1) / (kBlk / 64))),
                           dim3(kBlk), 0, h->stream, h->N, S, h->ooff.as<int>(),
                           h->doffs.as<int>(), h->dmask.as<unsigned>(), h->syma.as<int>(),
                           h->symb.as<int>(), h->dval.as<double>(), h->dsv.as<double>(),
                           flag.as<int>());
        GIS_HIP(h, hipGetLastError());
        int bad = 1;
        GIS_HIP(h, hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));
        h->sym_values = bad == 0;
    }
    if (h->amg) {
        h->amg->values_ok = false;
        h->amg->dic_values = false;
    }
    h->has_values = true;
    h->has_val32 = false;
    h->has_valT = false;
    h->has_val32T = false;
    return kOK;
}

int ensure_val32(gis_context* h) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    if (h->has_val32) return kOK;
    GIS_TRY(gather_both<float>(h, false, h->val32, h->sval32, h->dval32));
    h->has_val32 = true;
    return kOK;
}

// A^T of an LDU matrix: same pattern, lower <-> upper (coupling entries mirrored)
// float(A^T): the mixed-precision transpose solves of BiCG
int ensure_val32T(gis_context* h) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    if (h->has_val32T) return kOK;
    GIS_TRY(gather_both<float>(h, true, h->val32T, h->sval32T, h->dval32T));
    h->has_val32T = true;
    return kOK;
}

int ensure_valT(gis_context* h) {
    if (!h->has_values) return fail(h, kErrState, "values not set");
    if (h->has_valT) return kOK;
    GIS_TRY(gather_both<double>(h, true, h->valT, h->svalT, h->dvalT));
    h->has_valT = true;
    return kOK;
}

}  // namespace gis
