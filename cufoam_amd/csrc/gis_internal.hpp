// Host-side context of the gis C-ABI: one handle = one device, one HIP stream,
// one cached LDU pattern (device CSR + value permutation) and its workspaces.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "gis_device.hpp"

struct gis_context;

namespace gis {

// Owning device allocation that only grows.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        if (n == 0) n = 256;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
    ~DevMem() { release(); }
};

// SpMV schedule: how a tile's products are formed.  Every schedule yields
// bit-identical results (same per-row order), so the choice is pure tuning.
enum class SpmvKind : int {
    Sell = 0, Staged2K = 1, Staged4K = 2, Direct = 3, Sell2 = 4, SellC = 5, SellD = 6
};

struct Schedule {
    SpmvKind spmv64 = SpmvKind::Sell;       // FP64 kernels
    SpmvKind spmv32 = SpmvKind::Sell;       // FP32 kernels
    int block_iters = 64;                   // max iterations launched between host checks
    // CG direction p_k = r + beta p_{k-1}: 0 formed on the fly inside the SpMV
    // kernel's gathers; 1 written by a streaming pass first, so that the SpMV
    // gathers one vector.  Same arithmetic, bit-identical iterates.
    int fp64_direction = 0;
    int fp32_direction = 0;
    // tiles per workgroup of the CG iteration kernels (1, 2 or 4); partials stay
    // per tile, so the reduction tree and the results do not change
    int coarsen64 = 1;
    int coarsen32 = 1;
};

// Opt-in kernel timing with HIP events on the launch stream.  Slot 0: the
// SpMV-carrying kernel of each iteration; 1: the vector-update kernel of each
// iteration; 2: FP64 update kernel of the mixed solver; 3: standalone SpMV;
// 4: the separate direction pass.
constexpr int kTimerSlots = 5;

struct Timer {
    unsigned mask = 0;                      // bit s: time slot s
    bool open = false;                      // a start event awaits its end
    std::vector<hipEvent_t> ev;             // pairs (start, end)
    std::vector<int> tag;                   // slot of each pair
    int used = 0;                           // pairs recorded since the last collect
    double total_ms[kTimerSlots] = {};
    long long count[kTimerSlots] = {};
    int period = 1;                         // time one launch in `period` of each slot
    long long seen[kTimerSlots] = {};       // launches of each slot since gis_set_timing
};

// Batched systems: block-diagonal stack, each system padded to whole tiles.
struct BatchInfo {
    bool ready = false;
    int S = 0;
    std::vector<int> n, m, pad_off, pack_off, face_off, tile0;
    DevMem red_mem, meta, st, slot_mem;
    size_t slot_n = 0;                   // doubles in slot_mem (every one holds kSlotEmpty)
    bool slots_dirty = false;            // a reduction timed out: refill before the next solve
    RedWs* d_red = nullptr;
    int* d_tile_sys = nullptr;
    int* d_tile0 = nullptr;
    int* d_pad_off = nullptr;
    int* d_pack_off = nullptr;
    KState* st_host = nullptr;           // pinned
    hipStream_t cap_stream = nullptr;    // graph capture / replay stream
    hipGraphExec_t graph_exec = nullptr;
    bool graph_valid = false;
    int graph_iters = 0;
    void* graph_key = nullptr;
    bool graph_selld = false;            // the SpMV layout the captured graph uses
    bool graph_split = false;            // the direction form the captured graph uses
};

// GAMG hierarchy (preconditioned CG).  Level 0 is the handle's own matrix;
// every coarser level is a context of its own (pattern + values) on the
// handle's stream.
struct AmgLevel {
    gis_context* ctx = nullptr;   // levels >= 1
    int n = 0, m = 0, nc = 0;     // cells, faces, cells of the next level
    // transfer to the next level: cell -> aggregate; aggregate -> its cells and
    // the faces inside it; next-level face -> fine faces (2 f + flip); all ascending
    DevMem agg, mptr, mlist, iptr, ilist, fptr, flist;
    DevMem sm;                    // omega / diag
    DevMem vecs;                  // z, z2, t, r (4 n)
    DevMem sm32, vecs32;          // the same for the FP32 cycle (GAMG32)
};

struct AmgData {
    bool built = false;           // hierarchy of the current pattern
    bool values_ok = false;       // coarse values of the current values
    bool values32_ok = false;     // FP32 copies (GAMG32) of the current values
    double omega = 0.8;           // Jacobi damping
    int nu = 1;                   // sweeps before and after the coarse correction
    double cscale = 1.8;          // coarse-correction factor
    std::deque<AmgLevel> lv;
    DevMem ainv, dense_work;      // dense inverse of the last level (dense_n > 0)
    int dense_n = 0;
    DevMem ainv32;                // (float) ainv
    DevMem tail, tail32;          // per-level pointers of the single-workgroup tail
    int tail_start = 0;           // first tail level (0: none)
    int tail_levels = 0;
    // DIC: cells grouped by dependency level of the forward (lower faces) and
    // backward (upper faces) sweeps, and the reciprocal factored diagonal rD
    bool dic_levels = false;      // level lists of the current pattern
    bool dic_values = false;      // rD of the current values
    std::vector<int> dic_ptr_f, dic_ptr_b;   // level offsets (host: launch sizes)
    DevMem dic_cells_f, dic_cells_b, dic_rd;
    int dic_kind = 0;             // the factor in dic_rd: 4 DIC, 5 DILU
    void clear();
    ~AmgData();
};

}  // namespace gis

struct gis_context {
    int device = 0;
    double wall_tick = 0.0;     // seconds per tick of the device wall clock (wall_clock64)
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // pattern (cached per mesh)
    int N = 0;
    int M = 0;
    long long nnz = 0;
    bool has_pattern = false;
    bool has_values = false;
    bool has_val32 = false;
    bool has_valT = false;
    bool has_val32T = false;
    gis::DevMem rowptr, col, perm, nbr_cnt;     // CSR pattern, value permutation
    gis::DevMem lower_addr, upper_addr;         // LDU addressing copies
    gis::DevMem diag, lower, upper;             // LDU coefficients (device copies)
    // coupled interfaces: entry j adds A(cpl_row[j], cpl_col[j]) = cpl_val[j] after the
    // LDU entries of its row; cpl_partner[j] is the entry at the mirrored position
    int n_cpl = 0;
    gis::DevMem cpl_row, cpl_col, cpl_val, cpl_partner;
    gis::DevMem val, valT, val32;               // CSR values: A, A^T, float(A)
    long long n_slots = 0;                      // sliced layout (64-row slices)
    gis::DevMem soff, scol, sperm;              // slice offsets, slot columns / codes
    gis::DevMem sval, svalT, sval32;            // slot values: A, A^T, float(A)
    gis::DevMem val32T, sval32T;                // float(A^T): CSR and slot values
    // compressed slot columns: col = row + sbase[slot row] + sres[slot], int8
    // residuals (-128 marks padding); sflag[slice] = 1 where every residual fits
    gis::DevMem sres, sbase, sflag;
    // column-offset tables (SellD): per slice the sorted column offsets of its
    // rows (doffs from ooff[s]; ooff[s+1] == ooff[s]: no table, int32 columns),
    // a bit mask per row of the offsets it holds, the slot -> value code map
    // and the slot values (64 slots per table entry)
    gis::DevMem ooff, doffs, dmask, dmask8, dperm;
    long long n_dslots = 0;
    // SellD layout counts (gis_get_layout): slices, slices with a table, their
    // table entries, tabled slices with one-byte masks, sliced-row slots of the others
    long long lay_slices = 0, lay_table_slices = 0, lay_table_entries = 0;
    long long lay_byte_mask_slices = 0, lay_untabled_slots = 0;
    gis::DevMem dval, dvalT, dval32, dval32T;
    // symmetric storage of the SellD layout (every slice tabled): the FP64 slot
    // values of the table entries with column offset >= 0 only (dsv), and per
    // table entry the value bases its slots read (syma, symb); built per pattern
    // (sym_layout), used while the values are bitwise symmetric (sym_values) and
    // the caller has not switched it off (sym_enable)
    gis::DevMem syma, symb, dsv;
    gis::DevMem symx;   // per table entry {offset, syma, symb, 0}: one 16-byte load a slot
    long long n_sym_slots = 0;
    bool sym_layout = false;
    bool sym_values = false;
    bool sym_enable = true;
    gis::DevMem dinv;                           // per-row helper (1/a_ii or sqrt(1/a_ii))
    gis::AmgData* amg = nullptr;                // GAMG hierarchy (built on first use)

    // workspaces
    gis::DevMem vec;                            // solver vectors
    gis::DevMem io;                             // staging for host <-> device vectors
    gis::DevMem red_mem;
    gis::RedWs red{};
    int red_rows = -1;
    int slot_fault = 0;                         // gis_set_slot_fault (failure-path tests)
    gis::DevMem state_mem;
    gis::KState* st = nullptr;                  // device state
    gis::KState* st_host = nullptr;             // pinned mirror
    gis::DevMem hist;
    gis::DevMem scratch;                        // pattern-build temporaries
    gis::DevMem mstate;                         // per-column states of multi-RHS solves
    gis::KState* mstate_host = nullptr;         // pinned mirror
    gis::DevMem io2;                            // staging for multi-RHS / adjoint data

    gis::Schedule sched;
    gis::Timer timer;
    gis::BatchInfo batch;
    std::string err;
};

namespace gis {

// status codes (mirrored in include/gis.h)
enum : int {
    kOK = 0,
    kErrArg = 1,
    kErrHip = 2,
    kErrShape = 3,
    kErrState = 4,
    kErrNoMem = 5,
    kBreakdown = 10,
    kNotConverged = 11,
    kNaN = 12,
};

// makes `dev` the current device for the scope
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

int fail(gis_context* h, int code, const std::string& msg);
int hip_fail(gis_context* h, hipError_t e, const char* what);

#define GIS_HIP(h, expr)                                        \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return ::gis::hip_fail((h), _e, #expr); \
    } while (0)

#define GIS_TRY(expr)             \
    do {                          \
        int _s = (expr);          \
        if (_s != ::gis::kOK) return _s; \
    } while (0)

__host__ __device__ inline int tiles_of(long long n) { return (int)((n + kTile - 1) / kTile); }

// kernel timing (gis_solvers.hip)
int timer_begin(gis_context* h, int slot);    // records the start event of a pair
int timer_end(gis_context* h);                // records the end event of the open pair
int timer_collect(gis_context* h);            // after a host sync: accumulate the pairs
void timer_release(gis_context* h);

// reductions / state
// workspace layout for `rows` rows at base (nullptr: sizes only); false when
// the tree would need more than kMaxLevels levels
bool reduction_layout(int rows, double* base, RedWs* out, size_t* n_doubles, size_t* n_slots);
int ensure_reduction(gis_context* h, int rows);
int reset_state(gis_context* h, KState* st);   // fresh state for dot / NormFactor
int fill_slots(gis_context* h, double* slot, size_t n);   // every slot to the empty pattern
int ensure_state(gis_context* h);
int read_state(gis_context* h);

// pattern + values (gis_pattern.hip)
int build_pattern(gis_context* h, int N, int M, const int* d_lower_addr, const int* d_upper_addr);
// symmetric storage in use: a layout built for it, values checked symmetric,
// not switched off
inline bool use_sym(const gis_context* h) {
    return h->sym_layout && h->sym_values && h->sym_enable;
}
int gather_values(gis_context* h);          // val (and valT / val32 when present) from diag/lower/upper
int ensure_val32(gis_context* h);
int ensure_valT(gis_context* h);
int ensure_val32T(gis_context* h);

// kernels (gis_solvers.hip)
int launch_spmv(gis_context* h, const double* x, double* y);
int launch_spmv_t(gis_context* h, const double* x, double* y);
int launch_spmv32(gis_context* h, const float* x, float* y);
int device_dot(gis_context* h, const double* a, const double* b, double* out);
int device_norm_factor(gis_context* h, const double* d_b, const double* d_x, double* nf);
// the two NormFactor launches on a given workspace and state (results in st->aux[1])
int launch_norm_factor(gis_context* h, const RedWs& red, KState* st, const double* d_b,
                       const double* d_x);
// allocation-free entry point (gis_noalloc.hip): per-device set-up, run by gis_create
int noalloc_setup(gis_context* h);

struct SolveArgs {
    int method;
    double error;
    long long max_steps;    // <= 0: unbounded
    double omega;
    int power_steps;
    double int_bound;
    int hist_cap;
    double rel_tol = 0.0;     // > 0: stop also at rel_tol * initial residual
    long long min_steps = 0;  // no stop on convergence before this many iterations
    int detail = 0;           // history: 0 residual norms, 1 elapsed seconds
};

// device_solve method ids past GIS_RICHARDSON: 8 + pc for preconditioned CG
// (pc 0 none, 1 diagonal, 2 GAMG, 3 GAMG32, 4 DIC), and BiCG preconditioned by
// the diagonal.
constexpr int kMethodPcg0 = 8;
constexpr int kMethodPbicgDiag = 13;
constexpr int kMethodPbicgDilu = 14;
// DIC / DILU (gis_amg.hip): the reciprocal factored diagonal for kind 4 (DIC)
// or 5 (DILU), and z = M^-1 r by the level-scheduled sweeps with the forward
// sweep's coefficients cf and the backward sweep's cb
int ilu_prepare(gis_context* h, int kind);
int ilu_apply(gis_context* h, const double* r, double* z, const double* cf, const double* cb);
// Preconditioned CG (gis_amg.hip), as device_solve methods kMethodPcg0 + pc.
int solve_pcg(gis_context* h, const SolveArgs& a, int pc, double* x, const double* b);
int amg_levels(gis_context* h, int* n_levels);
int amg_level(gis_context* h, int l, int* n, int* m, int* agg, int* la, int* ua, double* diag,
              double* lower, double* upper);

// x (in/out) and b are device vectors owned by the handle's io workspace.
int device_solve(gis_context* h, const SolveArgs& a, double* d_x, const double* d_b,
                 int* hist_len, double* norm0, double* norm_final);

// nrhs columns of length N, each contiguous (X[k*N + i]); transpose: solve A^T X = B.
int device_solve_multi(gis_context* h, int method, int nrhs, bool transpose, const SolveArgs& a,
                       double* X, const double* B, int* iters, double* norm0, double* normf,
                       int* status);
// Mixed-precision multi-RHS solve (restarted refinement): FP64 residual
// R = B - A X, FP32 inner CG/BiCG on A D = R from D = 0 until its recursive
// residual falls below max(inner_rel_tol ||R||, rel_tol ||b|| / 2), X += D in
// FP64; per column until
// ||b - A x||_2 <= rel_tol ||b||_2.  hist_len / iters / rel_res / status per column.
struct MultiMixedArgs {
    double rel_tol;
    double inner_rel_tol;
    long long max_outer;
    long long max_inner;
    double int_bound;
    int hist_cap;
};
int device_solve_multi_mixed(gis_context* h, int method, int nrhs, const MultiMixedArgs& a,
                             double* X, const double* B, double* hist, int* outer,
                             int* iters, double* rel_res, int* status);

// dL_k/dA = -lambda_k x^T on the pattern, written into LDU-shaped gradients.
int device_sddmm_fold(gis_context* h, int nrhs, const double* Lam, const double* x,
                      double* d_diag, double* d_lower, double* d_upper);

// batched systems (gis_batch.hip); addressing is per system (local cell ids),
// faces and coefficients concatenated system by system
int batch_set_pattern(gis_context* h, int S, const int* n_cells, const int* n_faces,
                      const int* lower_addr, const int* upper_addr);
int batch_set_values(gis_context* h, const double* diag, const double* lower,
                     const double* upper, bool dev);
int batch_solve(gis_context* h, double* x_packed, const double* b_packed, bool dev,
                const double* error, long long max_steps, int use_graph, int* iters,
                double* final_res, int* status);
void batch_release(gis_context* h);

struct MixedArgs {
    double rel_tol;
    double inner_rel_tol;
    long long max_iter;
    int mode;               // 0 reliable updates, 1 restarted refinement
    int hist_cap;
};

// schedule autotuner (gis_tune.hip); what: 1 FP64 kernels, 2 FP32 kernels
int device_autotune(gis_context* h, int what, int iterations, int* spmv, int* direction,
                    int* from_cache, double* best_ms);
// the schedule a new pattern starts with (no timing): a cached autotuner choice
// for (device, N, nnz), else SellD with the separate direction pass when every
// slice has an offset table, else sliced rows
void pattern_schedule(gis_context* h);
void autotune_forget();

int device_solve_mixed(gis_context* h, const MixedArgs& a, double* d_x, const double* d_b,
                       int* hist_len, int* iters, double* rel_res);

}  // namespace gis
