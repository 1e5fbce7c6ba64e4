// Host side of the GAMG set-up: one coarsening step of the pairwise
// agglomeration (two pairing passes over the face graph, weights
// |upper| + |lower|), the coarse faces in upper-triangular order, the Galerkin
// values of the next level (for the next step's weights) and the transfer lists
// the device kernels read.  Every result is a function of the input alone: the
// pairing passes run in cell order, and every parallel part either writes
// disjoint outputs or places items with atomic cursors and then sorts each
// segment, so the lists, the coarse faces and the floating-point sums (one
// fixed order per output: members ascending, then faces ascending) do not
// depend on the thread count.  Plain C++17 (no HIP), so it also builds for
// host-only checks.
#pragma once

#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>
#include <thread>
#include <utility>
#include <vector>

namespace gis {
namespace agglom {

// Allocator of the set-up's arrays: sizing a vector leaves its elements
// uninitialised (they are written by parallel loops, which also take the page
// faults in parallel), and blocks of 4 MiB or more are mapped with transparent
// huge pages requested.
template <class T>
struct Uninit {
    using value_type = T;
    static constexpr size_t kBig = size_t(4) << 20;
    Uninit() = default;
    template <class U>
    Uninit(const Uninit<U>&) {}
    static size_t mapped(size_t bytes) { return (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1); }
    T* allocate(size_t n) {
        const size_t b = n * sizeof(T);
        if (b >= kBig) {
            void* p = mmap(nullptr, mapped(b), PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS,
                           -1, 0);
            if (p == MAP_FAILED) throw std::bad_alloc();
            madvise(p, mapped(b), MADV_HUGEPAGE);
            return static_cast<T*>(p);
        }
        void* p = std::malloc(b ? b : 1);
        if (!p) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t n) {
        const size_t b = n * sizeof(T);
        if (b >= kBig)
            munmap(p, mapped(b));
        else
            std::free(p);
    }
    template <class U>
    void construct(U* p) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
    template <class U>
    bool operator==(const Uninit<U>&) const { return true; }
    template <class U>
    bool operator!=(const Uninit<U>&) const { return false; }
};

template <class T>
using hvec = std::vector<T, Uninit<T>>;

struct HostLevel {
    int n = 0, m = 0;
    hvec<int> la, ua;
    hvec<double> diag, lower, upper;
};

// transfer from a level to the next: cell -> aggregate; aggregate -> its cells
// (mptr/mlist) and the faces inside it (iptr/ilist); next-level face -> its fine
// faces as 2 f + flip (fptr/flist); all lists ascending
struct Transfer {
    hvec<int> agg, mptr, mlist, iptr, ilist, fptr, flist;
};

// worker threads for the set-up: OMP_NUM_THREADS when set (the process's CPU
// share), else the hardware count, at most 64
inline int max_threads() {
    static const int t = [] {
        int v = 0;
        if (const char* e = std::getenv("OMP_NUM_THREADS")) v = std::atoi(e);
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 64));
    }();
    return t;
}

// threads for `work` items: one below 2^16, else at most one per 2^15 items
inline int threads_for(long long work) {
    if (work < (1 << 16)) return 1;
    return (int)std::max(1LL, std::min<long long>(max_threads(), work >> 15));
}

// f(lo, hi) over [0, n) in T contiguous chunks
template <class F>
void par_for(long long n, int T, F f) {
    if (T <= 1 || n < 2) {
        if (n > 0) f(0LL, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    const long long step = (n + T - 1) / T;
    for (int t = 1; t < T; ++t) {
        const long long lo = t * step, hi = std::min(n, lo + step);
        if (lo < hi) th.emplace_back([=, &f] { f(lo, hi); });
    }
    f(0LL, std::min(n, step));
    for (auto& x : th) x.join();
}

inline int fetch_inc(int* p) { return __atomic_fetch_add(p, 1, __ATOMIC_RELAXED); }

// n copies of v, written in parallel
template <class T>
hvec<T> filled(long long n, T v, int T_) {
    hvec<T> a((size_t)n);
    par_for(n, T_, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) a[i] = v;
    });
    return a;
}

// ptr[0] = 0, ptr[k + 1] = ptr[k] + cnt[k] for k < n (ptr has n + 1 entries)
inline void excl_scan(const hvec<int>& cnt, hvec<int>& ptr, int T) {
    const long long n = (long long)cnt.size();
    ptr.resize((size_t)n + 1);
    ptr[0] = 0;
    if (T <= 1 || n < (1 << 16)) {
        for (long long k = 0; k < n; ++k) ptr[k + 1] = ptr[k] + cnt[k];
        return;
    }
    const long long step = (n + T - 1) / T;
    std::vector<long long> part(T + 1, 0);
    par_for(T, T, [&](long long t0, long long t1) {
        for (long long t = t0; t < t1; ++t) {
            long long s = 0;
            for (long long k = t * step; k < std::min(n, (t + 1) * step); ++k) s += cnt[k];
            part[t + 1] = s;
        }
    });
    for (int t = 0; t < T; ++t) part[t + 1] += part[t];
    par_for(T, T, [&](long long t0, long long t1) {
        for (long long t = t0; t < t1; ++t) {
            long long s = part[t];
            for (long long k = t * step; k < std::min(n, (t + 1) * step); ++k) {
                s += cnt[k];
                ptr[k + 1] = (int)s;
            }
        }
    });
}

// sort a[lo, hi) ascending, carrying b along (segments are short)
inline void sort_pairs(int* a, int* b, int lo, int hi) {
    if (hi - lo <= 32) {
        for (int i = lo + 1; i < hi; ++i) {
            const int ka = a[i], kb = b ? b[i] : 0;
            int j = i - 1;
            while (j >= lo && a[j] > ka) {
                a[j + 1] = a[j];
                if (b) b[j + 1] = b[j];
                --j;
            }
            a[j + 1] = ka;
            if (b) b[j + 1] = kb;
        }
        return;
    }
    std::vector<std::pair<int, int>> v;
    v.reserve(hi - lo);
    for (int i = lo; i < hi; ++i) v.emplace_back(a[i], b ? b[i] : 0);
    std::sort(v.begin(), v.end());
    for (int i = lo; i < hi; ++i) {
        a[i] = v[i - lo].first;
        if (b) b[i] = v[i - lo].second;
    }
}

// counting-sort lists: for key k, the items i (ascending) with key[i] == k;
// keys < 0 are skipped.  item must be increasing in i, so that ascending items
// are index order.
template <class K, class I>
void bucket(int nkeys, const K& key, const I& item, hvec<int>& ptr, hvec<int>& list) {
    const long long n = (long long)key.size();
    const int T = threads_for(n + nkeys);
    hvec<int> cnt = filled(nkeys, 0, T);
    par_for(n, T, [&](long long lo, long long hi) {
        for (long long i = lo; i < hi; ++i)
            if (key[i] >= 0) fetch_inc(&cnt[key[i]]);
    });
    excl_scan(cnt, ptr, T);
    list.resize(ptr[nkeys]);
    hvec<int> cur(ptr.begin(), ptr.end() - 1);
    par_for(n, T, [&](long long lo, long long hi) {
        for (long long i = lo; i < hi; ++i)
            if (key[i] >= 0) list[fetch_inc(&cur[key[i]])] = item[i];
    });
    par_for(nkeys, T, [&](long long lo, long long hi) {
        for (long long k = lo; k < hi; ++k) sort_pairs(list.data(), nullptr, ptr[k], ptr[k + 1]);
    });
}

// rows of the face graph: faces where the cell is neighbour (face order), then
// faces where it is owner (face order); nb[p] is the other cell of face[p].
// The neighbour parts are placed by atomic cursors and sorted per row; the
// owner parts too, unless the faces are in owner order (OpenFOAM's
// upper-triangular order), where each owner part is a run of faces.
inline void graph_rows(int n, int m, const int* la, const int* ua, hvec<int>& rp,
                       hvec<int>& face, hvec<int>& nb) {
    const int T = threads_for((long long)m + n);
    hvec<int> lo_cnt = filled(n, 0, T), cnt = filled(n, 0, T);
    int unsorted = 0;
    par_for(m, T, [&](long long f0, long long f1) {
        int u = 0;
        for (long long f = f0; f < f1; ++f) {
            fetch_inc(&lo_cnt[ua[f]]);
            fetch_inc(&cnt[ua[f]]);
            fetch_inc(&cnt[la[f]]);
            if (f > 0 && la[f] < la[f - 1]) u = 1;
        }
        if (u) __atomic_store_n(&unsorted, 1, __ATOMIC_RELAXED);
    });
    excl_scan(cnt, rp, T);
    face.resize(2 * (size_t)m);
    nb.resize(2 * (size_t)m);
    hvec<int> cur_lo(n), cur_hi;
    if (unsorted) cur_hi.resize(n);
    par_for(n, T, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; ++c) {
            cur_lo[c] = rp[c];
            if (unsorted) cur_hi[c] = rp[c] + lo_cnt[c];
        }
    });
    par_for(m, T, [&](long long f0, long long f1) {
        for (long long f = f0; f < f1; ++f) {
            const int p = fetch_inc(&cur_lo[ua[f]]);
            face[p] = (int)f;
            nb[p] = la[f];
            if (unsorted) {
                const int q = fetch_inc(&cur_hi[la[f]]);
                face[q] = (int)f;
                nb[q] = ua[f];
            }
        }
    });
    if (!unsorted) {
        // the owner part of cell c: its faces' run starts where the owner changes
        par_for(m, T, [&](long long f0, long long f1) {
            long long run = f0;
            while (run > 0 && la[run - 1] == la[f0]) --run;
            for (long long f = f0; f < f1; ++f) {
                if (f > 0 && la[f] != la[f - 1]) run = f;
                const int c = la[f];
                const int q = rp[c] + lo_cnt[c] + (int)(f - run);
                face[q] = (int)f;
                nb[q] = ua[f];
            }
        });
    }
    par_for(n, T, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; ++c) {
            const int mid = rp[c] + lo_cnt[c];
            sort_pairs(face.data(), nb.data(), rp[c], mid);
            if (unsorted) sort_pairs(face.data(), nb.data(), mid, rp[c + 1]);
        }
    });
}

// One pairing pass in cell order: an unpaired cell pairs with its unpaired
// neighbour of largest weight (the first in row order on ties); with none
// left it joins the aggregate of its neighbour of largest weight, or stays
// alone.  Returns the aggregate count.
inline int pair_pass(int n, int m, const int* la, const int* ua, const double* w,
                     hvec<int>& agg) {
    hvec<int> rp, face, nb;
    graph_rows(n, m, la, ua, rp, face, nb);
    // the weights in row order, gathered in parallel, so that the sequential
    // pass below reads only streams and the aggregate map
    const int T = threads_for(2 * (long long)m);
    hvec<double> wt(2 * (size_t)m);
    par_for(2 * (long long)m, T, [&](long long p0, long long p1) {
        for (long long p = p0; p < p1; ++p) wt[p] = w[face[p]];
    });
    agg = filled(n, -1, threads_for(n));
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int best = -1;
        double bw = -1.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (agg[nb[p]] < 0 && wt[p] > bw) {
                bw = wt[p];
                best = nb[p];
            }
        if (best >= 0) {
            agg[i] = agg[best] = nc++;
            continue;
        }
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (wt[p] > bw) {
                bw = wt[p];
                best = nb[p];
            }
        agg[i] = best >= 0 ? agg[best] : nc++;
    }
    return nc;
}

// Distinct aggregate pairs of the fine faces in upper-triangular order
// (sorted by owner, then neighbour) as cla / cua; fcode[f] = 2 F + flip, or
// -1 - C for a face inside aggregate C; fptr / flist: per coarse face its fine
// faces as 2 f + flip, ascending.  With fw, wsum[F] = sum of fw over the fine
// faces of F in face order (from +0.0).  Returns the coarse face count.
inline int coarse_faces(int m, const int* la, const int* ua, const int* agg, int nc,
                        hvec<int>& fcode, hvec<int>& cla, hvec<int>& cua, hvec<int>& fptr,
                        hvec<int>& flist, const double* fw = nullptr,
                        hvec<double>* wsum = nullptr) {
    const int T = threads_for((long long)m + nc);
    hvec<int> cnt = filled(nc, 0, T), start;
    par_for(m, T, [&](long long f0, long long f1) {
        for (long long f = f0; f < f1; ++f) {
            const int a = agg[la[f]], b = agg[ua[f]];
            if (a != b) fetch_inc(&cnt[std::min(a, b)]);
        }
    });
    excl_scan(cnt, start, T);
    // per crossing face: (other aggregate) << 32 | (2 f + flip), grouped by owner
    hvec<long long> ent(start[nc]);
    {
        hvec<int> cur(start.begin(), start.end() - 1);
        par_for(m, T, [&](long long f0, long long f1) {
            for (long long f = f0; f < f1; ++f) {
                const int a = agg[la[f]], b = agg[ua[f]];
                if (a != b)
                    ent[fetch_inc(&cur[std::min(a, b)])] =
                        ((long long)std::max(a, b) << 32) | (2 * f + (a > b ? 1 : 0));
            }
        });
    }
    // sort each owner's entries by (other aggregate, face); count the distinct pairs
    par_for(nc, T, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; ++c) {
            long long* e = ent.data();
            const int lo = start[c], hi = start[c + 1];
            if (hi - lo <= 32) {
                for (int i = lo + 1; i < hi; ++i) {
                    const long long k = e[i];
                    int j = i - 1;
                    while (j >= lo && e[j] > k) {
                        e[j + 1] = e[j];
                        --j;
                    }
                    e[j + 1] = k;
                }
            } else {
                std::sort(e + lo, e + hi);
            }
            int u = 0;
            for (int p = lo; p < hi; ++p)
                if (p == lo || (e[p] >> 32) != (e[p - 1] >> 32)) ++u;
            cnt[c] = u;
        }
    });
    hvec<int> ustart;
    excl_scan(cnt, ustart, T);
    const int mc = ustart[nc];
    cla.resize(mc);
    cua.resize(mc);
    fptr.resize((size_t)mc + 1);
    fptr[mc] = start[nc];
    flist.resize(start[nc]);
    fcode.resize(m);
    if (wsum) wsum->resize(mc);
    par_for(nc, T, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; ++c) {
            int g = ustart[c] - 1;
            for (int p = start[c]; p < start[c + 1]; ++p) {
                const long long e = ent[p];
                if (p == start[c] || (e >> 32) != (ent[p - 1] >> 32)) {
                    ++g;
                    cla[g] = (int)c;
                    cua[g] = (int)(e >> 32);
                    fptr[g] = p;
                    if (wsum) (*wsum)[g] = 0.0;
                }
                const int code = (int)(e & 0xffffffffLL);
                flist[p] = code;
                fcode[code >> 1] = 2 * g + (code & 1);
                if (wsum) (*wsum)[g] = (*wsum)[g] + fw[code >> 1];
            }
        }
    });
    par_for(m, T, [&](long long f0, long long f1) {
        for (long long f = f0; f < f1; ++f) {
            const int a = agg[la[f]];
            if (a == agg[ua[f]]) fcode[f] = -1 - a;
        }
    });
    return mc;
}

// Cell and inside-face lists of a cell -> aggregate map with nc aggregates.
inline void member_lists(int n, int m, int nc, const hvec<int>& fcode, Transfer& t) {
    const int T = threads_for((long long)n + m);
    hvec<int> cells(n), inside(m), faces(m);
    par_for(n, T, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) cells[i] = (int)i;
    });
    par_for(m, T, [&](long long f0, long long f1) {
        for (long long f = f0; f < f1; ++f) {
            inside[f] = fcode[f] < 0 ? -1 - fcode[f] : -1;
            faces[f] = (int)f;
        }
    });
    bucket(nc, t.agg, cells, t.mptr, t.mlist);
    bucket(nc, inside, faces, t.iptr, t.ilist);
}

// Galerkin values of the next level: per aggregate the diagonals of its cells
// (ascending), then upper and lower of each face inside it (ascending); per
// coarse face the upper and lower of its fine faces (ascending, swapped when
// the coarse face runs against the fine one).
inline void restrict_values(const HostLevel& F, const Transfer& t, HostLevel& C) {
    const int T = threads_for((long long)F.n + F.m);
    C.diag.resize(C.n);
    C.upper.resize(C.m);
    C.lower.resize(C.m);
    par_for(C.n, T, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; ++c) {
            double s = 0.0;
            for (int p = t.mptr[c]; p < t.mptr[c + 1]; ++p) s = s + F.diag[t.mlist[p]];
            for (int p = t.iptr[c]; p < t.iptr[c + 1]; ++p) {
                const int f = t.ilist[p];
                s = s + F.upper[f];
                s = s + F.lower[f];
            }
            C.diag[c] = s;
        }
    });
    par_for(C.m, T, [&](long long g0, long long g1) {
        for (long long g = g0; g < g1; ++g) {
            double su = 0.0, sl = 0.0;
            for (int p = t.fptr[g]; p < t.fptr[g + 1]; ++p) {
                const int f = t.flist[p] >> 1;
                const bool flip = t.flist[p] & 1;
                su = su + (flip ? F.lower[f] : F.upper[f]);
                sl = sl + (flip ? F.upper[f] : F.lower[f]);
            }
            C.upper[g] = su;
            C.lower[g] = sl;
        }
    });
}

// One coarsening step from `cur`: two pairing passes (the second on the graph
// of the first pass's aggregates, weights summed over the fine faces in face
// order).  Returns false, leaving next and t untouched, when the step keeps more
// than 90% of the cells (stagnation).
inline bool coarsen(const HostLevel& cur, HostLevel& next, Transfer& t) {
    const int T = threads_for((long long)cur.m + cur.n);
    hvec<double> w(cur.m);
    par_for(cur.m, T, [&](long long f0, long long f1) {
        for (long long f = f0; f < f1; ++f) w[f] = std::fabs(cur.upper[f]) + std::fabs(cur.lower[f]);
    });
    hvec<int> agg1, fc1, la1, ua1, fp1, fl1, agg2;
    hvec<double> w1;   // summed over the fine faces of each face, in face order
    const int n1 = pair_pass(cur.n, cur.m, cur.la.data(), cur.ua.data(), w.data(), agg1);
    const int m1 = coarse_faces(cur.m, cur.la.data(), cur.ua.data(), agg1.data(), n1, fc1, la1,
                                ua1, fp1, fl1, w.data(), &w1);
    const int n2 = pair_pass(n1, m1, la1.data(), ua1.data(), w1.data(), agg2);
    if ((long long)n2 * 10 > (long long)cur.n * 9) return false;
    t.agg.resize(cur.n);
    par_for(cur.n, T, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) t.agg[i] = agg2[agg1[i]];
    });
    hvec<int> fcode;
    next.n = n2;
    next.m = coarse_faces(cur.m, cur.la.data(), cur.ua.data(), t.agg.data(), n2, fcode, next.la,
                          next.ua, t.fptr, t.flist);
    member_lists(cur.n, cur.m, n2, fcode, t);
    restrict_values(cur, t, next);
    return true;
}

}  // namespace agglom
}  // namespace gis
