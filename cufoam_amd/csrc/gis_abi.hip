// extern "C" entry points of include/gis.h.
#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/gis.h"
#include "gis_internal.hpp"

namespace gis {

int fail(gis_context* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

int hip_fail(gis_context* h, hipError_t e, const char* what) {
    if (h) h->err = std::string(what) + ": " + hipGetErrorString(e);
    return kErrHip;
}

namespace {

int upload(gis_context* h, DevMem& dst, const void* src, size_t bytes) {
    GIS_HIP(h, dst.ensure(bytes));
    if (bytes) GIS_HIP(h, hipMemcpyAsync(dst.p, src, bytes, hipMemcpyHostToDevice, h->stream));
    return kOK;
}

int copy_d2d(gis_context* h, DevMem& dst, const void* src, size_t bytes) {
    GIS_HIP(h, dst.ensure(bytes));
    if (bytes) GIS_HIP(h, hipMemcpyAsync(dst.p, src, bytes, hipMemcpyDeviceToDevice, h->stream));
    return kOK;
}

int set_values_common(gis_context* h) {
    GIS_TRY(gather_values(h));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return kOK;
}

bool detail_ok(const gis_params* p) {
    return !p || p->detail_type == GIS_DETAIL_ERROR || p->detail_type == GIS_DETAIL_TIME;
}

SolveArgs to_args(int method, const gis_params* p, int hist_cap) {
    gis_params d;
    gis_default_params(&d);
    if (!p) p = &d;
    SolveArgs a;
    a.method = method;
    a.error = p->error;
    a.max_steps = p->max_steps;
    a.omega = p->omega;
    a.power_steps = p->power_steps;
    a.int_bound = p->int_bound;
    a.hist_cap = hist_cap;
    a.detail = p->detail_type;
    return a;
}

// x and b are staged into the handle's io buffer: [x | b]
int solve_staged(gis_context* h, const SolveArgs& a, double* x_host_or_dev,
                 const double* b_host_or_dev, bool on_device, double* hist, int hist_cap,
                 int* hist_len, double* norm0, double* normf) {
    const size_t n = (size_t)h->N;
    GIS_HIP(h, h->io.ensure(sizeof(double) * 2 * n));
    double* dx = h->io.as<double>();
    double* db = dx + n;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(dx, x_host_or_dev, sizeof(double) * n, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(db, b_host_or_dev, sizeof(double) * n, kin, h->stream));
    int len = 0;
    const int st = device_solve(h, a, dx, db, &len, norm0, normf);
    if (st != kOK && st < kBreakdown) return st;
    GIS_HIP(h, hipMemcpyAsync(x_host_or_dev, dx, sizeof(double) * n, kout, h->stream));
    if (hist && hist_cap > 0) {
        const int m = std::min(len, std::min(hist_cap, a.hist_cap));
        GIS_HIP(h, hipMemcpyAsync(hist, h->hist.p, sizeof(double) * (size_t)m,
                                  hipMemcpyDeviceToHost, h->stream));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (hist_len) *hist_len = len;
    return st;
}

}  // namespace
}  // namespace gis

using namespace gis;

extern "C" {

GIS_API void gis_default_params(gis_params* p) {
    if (!p) return;
    p->error = 1e-4;
    p->max_steps = 0;
    p->omega = 0.0;
    p->power_steps = 30;
    p->int_bound = 1e-10;
    p->detail_type = GIS_DETAIL_ERROR;
}

GIS_API const char* gis_status_string(int s) {
    switch (s) {
        case GIS_OK: return "ok";
        case GIS_ERR_ARG: return "invalid argument";
        case GIS_ERR_HIP: return "HIP runtime error";
        case GIS_ERR_SHAPE: return "size mismatch or limit exceeded";
        case GIS_ERR_STATE: return "matrix pattern or values not set";
        case GIS_ERR_NOMEM: return "out of memory";
        case GIS_BREAKDOWN: return "BiCG breakdown";
        case GIS_NOT_CONVERGED: return "iteration cap reached";
        case GIS_NAN: return "residual is NaN";
        case GIS_ERR_SYNC: return "a device reduction did not complete";
        default: return "unknown status";
    }
}

GIS_API int gis_create(gis_handle* out, int device) {
    if (!out) return GIS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GIS_ERR_HIP;
    if (device < 0 || device >= ndev) return GIS_ERR_ARG;
    gis_context* h = new (std::nothrow) gis_context();
    if (!h) return GIS_ERR_NOMEM;
    h->device = device;
    DeviceGuard g(device);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return GIS_ERR_HIP;
    }
    h->own_stream = true;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
        wall_khz > 0)
        h->wall_tick = 1.0 / (1000.0 * wall_khz);
    if (ensure_state(h) != kOK || noalloc_setup(h) != kOK) {
        if (h->st_host) (void)hipHostFree(h->st_host);
        (void)hipStreamDestroy(h->stream);
        delete h;
        return GIS_ERR_HIP;
    }
    *out = h;
    return GIS_OK;
}

GIS_API int gis_destroy(gis_handle h) {
    if (!h) return GIS_ERR_ARG;
    {
        DeviceGuard g(h->device);
        (void)hipStreamSynchronize(h->stream);
        if (h->st_host) (void)hipHostFree(h->st_host);
        h->st_host = nullptr;
        if (h->mstate_host) (void)hipHostFree(h->mstate_host);
        h->mstate_host = nullptr;
        timer_release(h);
        batch_release(h);
        delete h->amg;
        h->amg = nullptr;
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        delete h;   // DevMem destructors free device memory with the device current
    }
    return GIS_OK;
}

GIS_API int gis_set_stream(gis_handle h, void* stream) {
    if (!h) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (h->own_stream) GIS_HIP(h, hipStreamDestroy(h->stream));
    h->own_stream = false;
    h->stream = static_cast<hipStream_t>(stream);
    return GIS_OK;
}

GIS_API int gis_reset_stream(gis_handle h) {
    if (!h) return GIS_ERR_ARG;
    if (h->own_stream) return GIS_OK;
    DeviceGuard g(h->device);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    GIS_HIP(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
    return GIS_OK;
}

GIS_API int gis_synchronize(gis_handle h) {
    if (!h) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

GIS_API const char* gis_last_error(gis_handle h) { return h ? h->err.c_str() : "null handle"; }

static int set_pattern_impl(gis_handle h, int n_cells, int n_faces, const int* la,
                            const int* ua, bool dev) {
    if (!h) return GIS_ERR_ARG;
    if (n_cells < 1 || n_faces < 0 || (n_faces > 0 && (!la || !ua)))
        return fail(h, GIS_ERR_ARG, "set_pattern: need n_cells >= 1, n_faces >= 0, addressing");
    DeviceGuard g(h->device);
    h->batch.ready = false;   // the handle now holds one system
    h->batch.graph_valid = false;
    h->n_cpl = 0;
    const size_t b = sizeof(int) * (size_t)n_faces;
    if (dev) {
        GIS_TRY(copy_d2d(h, h->lower_addr, la, b));
        GIS_TRY(copy_d2d(h, h->upper_addr, ua, b));
    } else {
        GIS_TRY(upload(h, h->lower_addr, la, b));
        GIS_TRY(upload(h, h->upper_addr, ua, b));
    }
    GIS_TRY(build_pattern(h, n_cells, n_faces, h->lower_addr.as<int>(), h->upper_addr.as<int>()));
    pattern_schedule(h);
    GIS_TRY(ensure_reduction(h, n_cells));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

GIS_API int gis_set_pattern(gis_handle h, int n_cells, int n_faces, const int* lower_addr,
                            const int* upper_addr) {
    return set_pattern_impl(h, n_cells, n_faces, lower_addr, upper_addr, false);
}

GIS_API int gis_set_pattern_device(gis_handle h, int n_cells, int n_faces,
                                   const int* d_lower_addr, const int* d_upper_addr) {
    return set_pattern_impl(h, n_cells, n_faces, d_lower_addr, d_upper_addr, true);
}

static int set_values_impl(gis_handle h, const double* diag, const double* lower,
                           const double* upper, bool dev) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "set_values before set_pattern");
    if (!diag || (h->M > 0 && (!lower || !upper))) return fail(h, GIS_ERR_ARG, "null coefficients");
    DeviceGuard g(h->device);
    const size_t bn = sizeof(double) * (size_t)h->N;
    const size_t bm = sizeof(double) * (size_t)h->M;
    if (dev) {
        GIS_TRY(copy_d2d(h, h->diag, diag, bn));
        GIS_TRY(copy_d2d(h, h->lower, lower, bm));
        GIS_TRY(copy_d2d(h, h->upper, upper, bm));
    } else {
        GIS_TRY(upload(h, h->diag, diag, bn));
        GIS_TRY(upload(h, h->lower, lower, bm));
        GIS_TRY(upload(h, h->upper, upper, bm));
    }
    return set_values_common(h);
}

GIS_API int gis_set_values(gis_handle h, const double* diag, const double* lower,
                           const double* upper) {
    return set_values_impl(h, diag, lower, upper, false);
}

GIS_API int gis_set_values_device(gis_handle h, const double* d_diag, const double* d_lower,
                                  const double* d_upper) {
    return set_values_impl(h, d_diag, d_lower, d_upper, true);
}

GIS_API int gis_set_interfaces(gis_handle h, int n_couples, const int* row, const int* col) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_pattern || h->batch.ready)
        return fail(h, GIS_ERR_STATE, "set_interfaces needs a single-system pattern");
    if (n_couples < 0 || (n_couples > 0 && (!row || !col)))
        return fail(h, GIS_ERR_ARG, "set_interfaces: need n_couples >= 0 and row/col arrays");
    const int N = h->N;
    for (int j = 0; j < n_couples; ++j)
        if (row[j] < 0 || row[j] >= N || col[j] < 0 || col[j] >= N || row[j] == col[j])
            return fail(h, GIS_ERR_ARG,
                        "interface entries need 0 <= row, col < n_cells and row != col");
    // mirrored entry of each (r, c): the k-th (r, c) in entry order pairs with the
    // k-th (c, r), as the two halves of a cyclic patch list them
    std::vector<long long> key(n_couples), mkey(n_couples);
    std::vector<int> a(n_couples), b(n_couples), partner(n_couples);
    for (int j = 0; j < n_couples; ++j) {
        key[j] = (long long)row[j] * N + col[j];
        mkey[j] = (long long)col[j] * N + row[j];
        a[j] = b[j] = j;
    }
    std::stable_sort(a.begin(), a.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::stable_sort(b.begin(), b.end(), [&](int x, int y) { return mkey[x] < mkey[y]; });
    for (int i = 0; i < n_couples; ++i) {
        if (key[a[i]] != mkey[b[i]])
            return fail(h, GIS_ERR_ARG,
                        "interface entries must be structurally symmetric: every (row, col) "
                        "needs a (col, row)");
        partner[a[i]] = b[i];
    }
    DeviceGuard g(h->device);
    const size_t bi = sizeof(int) * (size_t)n_couples;
    h->n_cpl = 0;
    if (n_couples > 0) {
        GIS_TRY(upload(h, h->cpl_row, row, bi));
        GIS_TRY(upload(h, h->cpl_col, col, bi));
        GIS_TRY(upload(h, h->cpl_partner, partner.data(), bi));
        GIS_HIP(h, h->cpl_val.ensure(sizeof(double) * (size_t)n_couples));
        GIS_HIP(h, hipMemsetAsync(h->cpl_val.p, 0, sizeof(double) * (size_t)n_couples,
                                  h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));   // partner is a host temporary
    }
    h->n_cpl = n_couples;
    GIS_TRY(build_pattern(h, N, h->M, h->lower_addr.as<int>(), h->upper_addr.as<int>()));
    pattern_schedule(h);
    GIS_TRY(ensure_reduction(h, N));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

static int set_coupled_values_impl(gis_handle h, const double* diag, const double* lower,
                                   const double* upper, const double* coeffs, bool dev) {
    if (!h) return GIS_ERR_ARG;
    if (h->n_cpl > 0 && !coeffs) return fail(h, GIS_ERR_ARG, "null interface coefficients");
    if (h->has_pattern && h->n_cpl > 0) {
        DeviceGuard g(h->device);
        const size_t bc = sizeof(double) * (size_t)h->n_cpl;
        if (dev) GIS_TRY(copy_d2d(h, h->cpl_val, coeffs, bc));
        else GIS_TRY(upload(h, h->cpl_val, coeffs, bc));
    }
    return set_values_impl(h, diag, lower, upper, dev);
}

GIS_API int gis_set_coupled_values(gis_handle h, const double* diag, const double* lower,
                                   const double* upper, const double* coeffs) {
    return set_coupled_values_impl(h, diag, lower, upper, coeffs, false);
}

GIS_API int gis_set_coupled_values_device(gis_handle h, const double* d_diag,
                                          const double* d_lower, const double* d_upper,
                                          const double* d_coeffs) {
    return set_coupled_values_impl(h, d_diag, d_lower, d_upper, d_coeffs, true);
}

GIS_API int gis_get_size(gis_handle h, int* n_rows, long long* nnz) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "pattern not set");
    if (n_rows) *n_rows = h->N;
    if (nnz) *nnz = h->nnz;
    return GIS_OK;
}

GIS_API int gis_get_layout(gis_handle h, gis_layout* out) {
    if (!h || !out) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "pattern not set");
    out->n_rows = h->N;
    out->nnz = h->nnz;
    out->slices = h->lay_slices;
    out->sell_slots = h->n_slots;
    out->table_slices = h->lay_table_slices;
    out->table_entries = h->lay_table_entries;
    out->byte_mask_slices = h->lay_byte_mask_slices;
    out->untabled_slots = h->lay_untabled_slots;
    return GIS_OK;
}

static int get_csr_impl(gis_handle h, int* rp, int* ci, double* v, bool transpose) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "pattern not set");
    DeviceGuard g(h->device);
    if (rp)
        GIS_HIP(h, hipMemcpyAsync(rp, h->rowptr.p, sizeof(int) * ((size_t)h->N + 1),
                                  hipMemcpyDeviceToHost, h->stream));
    if (ci)
        GIS_HIP(h, hipMemcpyAsync(ci, h->col.p, sizeof(int) * (size_t)h->nnz,
                                  hipMemcpyDeviceToHost, h->stream));
    if (v) {
        if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
        if (transpose) GIS_TRY(ensure_valT(h));
        GIS_HIP(h, hipMemcpyAsync(v, transpose ? h->valT.p : h->val.p,
                                  sizeof(double) * (size_t)h->nnz, hipMemcpyDeviceToHost,
                                  h->stream));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

GIS_API int gis_get_csr(gis_handle h, int* row_offsets, int* column_indices, double* values) {
    return get_csr_impl(h, row_offsets, column_indices, values, false);
}

GIS_API int gis_get_csr_transpose(gis_handle h, int* row_offsets, int* column_indices,
                                  double* values) {
    return get_csr_impl(h, row_offsets, column_indices, values, true);
}

GIS_API int gis_spmv_device(gis_handle h, const double* d_x, double* d_y) {
    if (!h || !d_x || !d_y) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_TRY(launch_spmv(h, d_x, d_y));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

GIS_API int gis_spmv_device_async(gis_handle h, const double* d_x, double* d_y) {
    if (!h || !d_x || !d_y) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_TRY(timer_begin(h, 3));
    GIS_TRY(launch_spmv(h, d_x, d_y));
    return timer_end(h);
}

GIS_API int gis_spmv32_device_async(gis_handle h, const float* d_x, float* d_y) {
    if (!h || !d_x || !d_y) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_TRY(timer_begin(h, 3));
    GIS_TRY(launch_spmv32(h, d_x, d_y));
    return timer_end(h);
}

GIS_API int gis_set_timing(gis_handle h, int slot_mask) {
    if (!h) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    GIS_TRY(timer_collect(h));
    h->timer.mask = (unsigned)slot_mask & ((1u << kTimerSlots) - 1);
    for (int i = 0; i < kTimerSlots; ++i) {
        h->timer.total_ms[i] = 0.0;
        h->timer.count[i] = 0;
        h->timer.seen[i] = 0;
    }
    return GIS_OK;
}

GIS_API int gis_set_symmetric_storage(gis_handle h, int enable) {
    if (!h || (enable != 0 && enable != 1)) return GIS_ERR_ARG;
    h->sym_enable = enable != 0;
    if (h->amg)   // the GAMG levels follow the handle
        for (auto& L : h->amg->lv)
            if (L.ctx) L.ctx->sym_enable = h->sym_enable;
    return GIS_OK;
}

GIS_API int gis_get_symmetric_storage(gis_handle h, int* active, long long* stored_slots) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "pattern not set");
    if (active) *active = use_sym(h) ? 1 : 0;
    if (stored_slots) *stored_slots = h->sym_layout ? h->n_sym_slots : 0;
    return GIS_OK;
}

GIS_API int gis_set_timing_sample(gis_handle h, int period) {
    if (!h) return GIS_ERR_ARG;
    if (period < 1) return fail(h, GIS_ERR_ARG, "period must be >= 1");
    h->timer.period = period;
    return GIS_OK;
}

GIS_API int gis_get_timing(gis_handle h, int slot, double* total_ms, long long* count) {
    if (!h || slot < 0 || slot >= kTimerSlots) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    GIS_TRY(timer_collect(h));
    if (total_ms) *total_ms = h->timer.total_ms[slot];
    if (count) *count = h->timer.count[slot];
    return GIS_OK;
}

GIS_API int gis_spmv_transpose_device(gis_handle h, const double* d_x, double* d_y) {
    if (!h || !d_x || !d_y) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    GIS_TRY(launch_spmv_t(h, d_x, d_y));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return GIS_OK;
}

GIS_API int gis_dot_device(gis_handle h, const double* d_a, const double* d_b, double* out) {
    if (!h || !d_a || !d_b || !out) return GIS_ERR_ARG;
    if (!h->has_pattern) return fail(h, GIS_ERR_STATE, "pattern not set (vector length)");
    DeviceGuard g(h->device);
    return device_dot(h, d_a, d_b, out);
}

GIS_API int gis_set_schedule(gis_handle h, const gis_schedule* s) {
    if (!h || !s) return GIS_ERR_ARG;
    auto coarsen_ok = [](int c) { return c <= 0 || c == 1 || c == 2 || c == 4; };
    if (s->spmv64 > 6 || s->spmv32 > 6 || s->fp64_direction > 1 || s->fp32_direction > 1 ||
        !coarsen_ok(s->coarsen64) || !coarsen_ok(s->coarsen32))
        return fail(h, GIS_ERR_ARG, "schedule field out of range");
    if (s->spmv64 >= 0) h->sched.spmv64 = static_cast<SpmvKind>(s->spmv64);
    if (s->spmv32 >= 0) h->sched.spmv32 = static_cast<SpmvKind>(s->spmv32);
    if (s->block_iters > 0) h->sched.block_iters = s->block_iters;
    if (s->fp64_direction >= 0) h->sched.fp64_direction = s->fp64_direction;
    if (s->fp32_direction >= 0) h->sched.fp32_direction = s->fp32_direction;
    if (s->coarsen64 > 0) h->sched.coarsen64 = s->coarsen64;
    if (s->coarsen32 > 0) h->sched.coarsen32 = s->coarsen32;
    return GIS_OK;
}

GIS_API int gis_get_schedule(gis_handle h, gis_schedule* s) {
    if (!h || !s) return GIS_ERR_ARG;
    s->spmv64 = static_cast<int>(h->sched.spmv64);
    s->spmv32 = static_cast<int>(h->sched.spmv32);
    s->block_iters = h->sched.block_iters;
    s->fp64_direction = h->sched.fp64_direction;
    s->fp32_direction = h->sched.fp32_direction;
    s->coarsen64 = h->sched.coarsen64;
    s->coarsen32 = h->sched.coarsen32;
    return GIS_OK;
}

GIS_API int gis_autotune(gis_handle h, int targets, int iterations, gis_schedule* chosen,
                         int* from_cache, double* best_ms) {
    if (!h) return GIS_ERR_ARG;
    if (targets <= 0 || (targets & ~(GIS_TUNE_FP64 | GIS_TUNE_FP32)))
        return fail(h, GIS_ERR_ARG, "autotune targets must be GIS_TUNE_FP64 and/or GIS_TUNE_FP32");
    DeviceGuard g(h->device);
    int cached_all = 1;
    for (int what = 1; what <= 2; ++what) {
        if (!(targets & what)) continue;
        int spmv = 0, dir = 0, cached = 0;
        double ms = 0.0;
        GIS_TRY(device_autotune(h, what, iterations, &spmv, &dir, &cached, &ms));
        cached_all &= cached;
        if (best_ms) best_ms[what - 1] = ms;
    }
    if (from_cache) *from_cache = cached_all;
    if (chosen) return gis_get_schedule(h, chosen);
    return GIS_OK;
}

GIS_API int gis_autotune_clear(void) {
    autotune_forget();
    return GIS_OK;
}

static int norm_factor_impl(gis_handle h, const double* b, const double* x, double* nf,
                            bool dev) {
    if (!h || !b || !x || !nf) return GIS_ERR_ARG;
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    const size_t n = (size_t)h->N;
    GIS_HIP(h, h->io.ensure(sizeof(double) * 2 * n));
    double* dx = h->io.as<double>();
    double* db = dx + n;
    const hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    GIS_HIP(h, hipMemcpyAsync(dx, x, sizeof(double) * n, k, h->stream));
    GIS_HIP(h, hipMemcpyAsync(db, b, sizeof(double) * n, k, h->stream));
    return device_norm_factor(h, db, dx, nf);
}

GIS_API int gis_norm_factor(gis_handle h, const double* b, const double* x, double* nf) {
    return norm_factor_impl(h, b, x, nf, false);
}

GIS_API int gis_norm_factor_device(gis_handle h, const double* d_b, const double* d_x,
                                   double* nf) {
    return norm_factor_impl(h, d_b, d_x, nf, true);
}

static int solve_impl(gis_handle h, int method, double* x, const double* b,
                      const gis_params* params, double* hist, int hist_cap, int* hist_len,
                      bool dev) {
    if (!h || !x || !b) return GIS_ERR_ARG;
    if (method < GIS_CG || method > GIS_RICHARDSON) return fail(h, GIS_ERR_ARG, "unknown method");
    if (!detail_ok(params)) return fail(h, GIS_ERR_ARG, "detail_type must be ERROR or TIME");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    SolveArgs a = to_args(method, params, std::max(1, hist_cap));
    return solve_staged(h, a, x, b, dev, hist, hist_cap, hist_len, nullptr, nullptr);
}

GIS_API int gis_solve(gis_handle h, int method, double* x, const double* b,
                      const gis_params* params, double* hist, int hist_cap, int* hist_len) {
    return solve_impl(h, method, x, b, params, hist, hist_cap, hist_len, false);
}

GIS_API int gis_solve_device(gis_handle h, int method, double* d_x, const double* d_b,
                             const gis_params* params, double* hist, int hist_cap,
                             int* hist_len) {
    return solve_impl(h, method, d_x, d_b, params, hist, hist_cap, hist_len, true);
}

static int pcg_impl(gis_handle h, int precond, double* x, const double* b,
                    const gis_params* params, double* hist, int hist_cap, int* hist_len,
                    bool dev) {
    if (!h || !x || !b) return GIS_ERR_ARG;
    if (precond < GIS_PRECOND_NONE || precond > GIS_PRECOND_DIC)
        return fail(h, GIS_ERR_ARG, "unknown preconditioner");
    if (!detail_ok(params)) return fail(h, GIS_ERR_ARG, "detail_type must be ERROR or TIME");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    SolveArgs a = to_args(kMethodPcg0 + precond, params, std::max(1, hist_cap));
    return solve_staged(h, a, x, b, dev, hist, hist_cap, hist_len, nullptr, nullptr);
}

// GIS_PRECOND_NONE: BiCG itself; GIS_PRECOND_DIAGONAL: BiCG preconditioned by D;
// GIS_PRECOND_DILU: by OpenFOAM's DILU
static int pbicg_method(gis_handle h, int precond, int* method) {
    if (precond == GIS_PRECOND_NONE) {
        *method = GIS_BICG;
        return GIS_OK;
    }
    if (precond == GIS_PRECOND_DIAGONAL) {
        *method = kMethodPbicgDiag;
        return GIS_OK;
    }
    if (precond == GIS_PRECOND_DILU) {
        *method = kMethodPbicgDilu;
        return GIS_OK;
    }
    return fail(h, GIS_ERR_ARG, "PBiCG takes the none, diagonal or DILU preconditioner");
}

static int pbicg_impl(gis_handle h, int precond, double* x, const double* b,
                      const gis_params* params, double* hist, int hist_cap, int* hist_len,
                      bool dev) {
    if (!h || !x || !b) return GIS_ERR_ARG;
    int method = 0;
    GIS_TRY(pbicg_method(h, precond, &method));
    if (!detail_ok(params)) return fail(h, GIS_ERR_ARG, "detail_type must be ERROR or TIME");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    SolveArgs a = to_args(method, params, std::max(1, hist_cap));
    return solve_staged(h, a, x, b, dev, hist, hist_cap, hist_len, nullptr, nullptr);
}

GIS_API int gis_solve_pbicg(gis_handle h, int precond, double* x, const double* b,
                            const gis_params* params, double* hist, int hist_cap,
                            int* hist_len) {
    return pbicg_impl(h, precond, x, b, params, hist, hist_cap, hist_len, false);
}

GIS_API int gis_solve_pbicg_device(gis_handle h, int precond, double* d_x, const double* d_b,
                                   const gis_params* params, double* hist, int hist_cap,
                                   int* hist_len) {
    return pbicg_impl(h, precond, d_x, d_b, params, hist, hist_cap, hist_len, true);
}

GIS_API int gis_solve_pcg(gis_handle h, int precond, double* x, const double* b,
                          const gis_params* params, double* hist, int hist_cap, int* hist_len) {
    return pcg_impl(h, precond, x, b, params, hist, hist_cap, hist_len, false);
}

GIS_API int gis_solve_pcg_device(gis_handle h, int precond, double* d_x, const double* d_b,
                                 const gis_params* params, double* hist, int hist_cap,
                                 int* hist_len) {
    return pcg_impl(h, precond, d_x, d_b, params, hist, hist_cap, hist_len, true);
}

GIS_API int gis_set_slot_fault(gis_handle h, int polls) {
    if (!h) return GIS_ERR_ARG;
    if (polls < 0) return fail(h, GIS_ERR_ARG, "polls must be >= 0");
    h->slot_fault = polls;
    if (h->st) {   // the live state too, for dot and NormFactor between solves
        h->st_host->slot_fault = polls;
        GIS_HIP(h, hipMemcpyAsync(&h->st->slot_fault, &h->st_host->slot_fault, sizeof(int),
                                  hipMemcpyHostToDevice, h->stream));
        GIS_HIP(h, hipStreamSynchronize(h->stream));
    }
    return GIS_OK;
}

GIS_API int gis_set_gamg(gis_handle h, double omega, int nu, double cscale) {
    if (!h) return GIS_ERR_ARG;
    if (!(omega > 0.0 && omega < 2.0) || nu < 1 || nu > 8 || !(cscale > 0.0 && cscale < 2.0))
        return fail(h, GIS_ERR_ARG, "GAMG needs 0 < omega < 2, 1 <= nu <= 8, 0 < cscale < 2");
    if (!h->amg) {
        h->amg = new (std::nothrow) gis::AmgData();
        if (!h->amg) return GIS_ERR_NOMEM;
    }
    h->amg->omega = omega;
    h->amg->nu = nu;
    h->amg->cscale = cscale;
    h->amg->values_ok = false;
    return GIS_OK;
}

GIS_API int gis_gamg_levels(gis_handle h, int* n_levels) {
    if (!h || !n_levels) return GIS_ERR_ARG;
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    return amg_levels(h, n_levels);
}

GIS_API int gis_gamg_level(gis_handle h, int level, int* n_cells, int* n_faces, int* agg,
                           int* lower_addr, int* upper_addr, double* diag, double* lower,
                           double* upper) {
    if (!h) return GIS_ERR_ARG;
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    return amg_level(h, level, n_cells, n_faces, agg, lower_addr, upper_addr, diag, lower, upper);
}

static int solve_mixed_impl(gis_handle h, double* x, const double* b, double rel_tol,
                            double inner_rel_tol, long long max_iter, int mode, double* hist,
                            int hist_cap, int* hist_len, int* iterations, double* rel_residual,
                            bool dev) {
    if (!h || !x || !b) return GIS_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(h, GIS_ERR_ARG, "mixed mode must be 0 or 1");
    if (!(rel_tol >= 0.0) || !(inner_rel_tol >= 0.0))
        return fail(h, GIS_ERR_ARG, "tolerances must be >= 0");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    const size_t n = (size_t)h->N;
    GIS_HIP(h, h->io.ensure(sizeof(double) * 2 * n));
    double* dx = h->io.as<double>();
    double* db = dx + n;
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(dx, x, sizeof(double) * n, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(db, b, sizeof(double) * n, kin, h->stream));
    MixedArgs a{rel_tol, inner_rel_tol, max_iter, mode, std::max(1, hist_cap)};
    int len = 0, it = 0;
    double rr = 0.0;
    const int st = device_solve_mixed(h, a, dx, db, &len, &it, &rr);
    if (st != GIS_OK && st < GIS_BREAKDOWN) return st;
    GIS_HIP(h, hipMemcpyAsync(x, dx, sizeof(double) * n, kout, h->stream));
    if (hist && hist_cap > 0) {
        const int m = std::min(len, hist_cap);
        GIS_HIP(h, hipMemcpyAsync(hist, h->hist.p, sizeof(double) * (size_t)m,
                                  hipMemcpyDeviceToHost, h->stream));
    }
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (hist_len) *hist_len = len;
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = rr;
    return st;
}

GIS_API int gis_solve_mixed(gis_handle h, double* x, const double* b, double rel_tol,
                            double inner_rel_tol, long long max_iter, int mode, double* hist,
                            int hist_cap, int* hist_len, int* iterations, double* rel_residual) {
    return solve_mixed_impl(h, x, b, rel_tol, inner_rel_tol, max_iter, mode, hist, hist_cap,
                            hist_len, iterations, rel_residual, false);
}

GIS_API int gis_solve_mixed_device(gis_handle h, double* d_x, const double* d_b,
                                   double rel_tol, double inner_rel_tol, long long max_iter,
                                   int mode, double* hist, int hist_cap, int* hist_len,
                                   int* iterations, double* rel_residual) {
    return solve_mixed_impl(h, d_x, d_b, rel_tol, inner_rel_tol, max_iter, mode, hist,
                            hist_cap, hist_len, iterations, rel_residual, true);
}

static int solve_multi_impl(gis_handle h, int method, int nrhs, double* X, const double* B,
                            const gis_params* params, int transpose, int* iterations,
                            double* final_residual, int* status, bool dev) {
    if (!h || !X || !B) return GIS_ERR_ARG;
    if (nrhs < 1 || nrhs > 8) return fail(h, GIS_ERR_ARG, "nrhs must be in 1..8");
    if (method != GIS_CG && method != GIS_BICG)
        return fail(h, GIS_ERR_ARG, "multi-RHS solves support GIS_CG and GIS_BICG");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    const size_t n = (size_t)h->N * nrhs;
    GIS_HIP(h, h->io2.ensure(sizeof(double) * 2 * n));
    double* dx = h->io2.as<double>();
    double* db = dx + n;
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(dx, X, sizeof(double) * n, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(db, B, sizeof(double) * n, kin, h->stream));
    SolveArgs a = to_args(method, params, 1);
    const int st = device_solve_multi(h, method, nrhs, transpose != 0, a, dx, db, iterations,
                                      nullptr, final_residual, status);
    if (st != GIS_OK && st < GIS_BREAKDOWN) return st;
    GIS_HIP(h, hipMemcpyAsync(X, dx, sizeof(double) * n, kout, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return st;
}

GIS_API int gis_solve_multi(gis_handle h, int method, int nrhs, double* X, const double* B,
                            const gis_params* params, int transpose, int* iterations,
                            double* final_residual, int* status) {
    return solve_multi_impl(h, method, nrhs, X, B, params, transpose, iterations,
                            final_residual, status, false);
}

GIS_API int gis_solve_multi_device(gis_handle h, int method, int nrhs, double* d_X,
                                   const double* d_B, const gis_params* params, int transpose,
                                   int* iterations, double* final_residual, int* status) {
    return solve_multi_impl(h, method, nrhs, d_X, d_B, params, transpose, iterations,
                            final_residual, status, true);
}

static int solve_multi_mixed_impl(gis_handle h, int method, int nrhs, double* X,
                                  const double* B, double rel_tol, double inner_rel_tol,
                                  long long max_outer, long long max_inner, double* hist,
                                  int hist_cap, int* outer_steps, int* iterations,
                                  double* rel_residual, int* status, bool dev) {
    if (!h || !X || !B) return GIS_ERR_ARG;
    if (nrhs < 1 || nrhs > 8) return fail(h, GIS_ERR_ARG, "nrhs must be in 1..8");
    if (method != GIS_CG && method != GIS_BICG)
        return fail(h, GIS_ERR_ARG, "mixed multi-RHS solves support GIS_CG and GIS_BICG");
    if (!(rel_tol >= 0.0) || !(inner_rel_tol >= 0.0) || hist_cap < 0)
        return fail(h, GIS_ERR_ARG, "tolerances and hist_cap must be non-negative");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    const size_t n = (size_t)h->N * nrhs;
    GIS_HIP(h, h->io2.ensure(sizeof(double) * 2 * n));
    double* dx = h->io2.as<double>();
    double* db = dx + n;
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(dx, X, sizeof(double) * n, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(db, B, sizeof(double) * n, kin, h->stream));
    MultiMixedArgs a{};
    a.rel_tol = rel_tol;
    a.inner_rel_tol = inner_rel_tol;
    a.max_outer = max_outer;
    a.max_inner = max_inner;
    a.int_bound = 1e-10;
    a.hist_cap = hist ? hist_cap : 0;
    const int st = device_solve_multi_mixed(h, method, nrhs, a, dx, db, hist, outer_steps,
                                            iterations, rel_residual, status);
    if (st != GIS_OK && st < GIS_BREAKDOWN) return st;
    GIS_HIP(h, hipMemcpyAsync(X, dx, sizeof(double) * n, kout, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return st;
}

GIS_API int gis_solve_multi_mixed(gis_handle h, int method, int nrhs, double* X,
                                  const double* B, double rel_tol, double inner_rel_tol,
                                  long long max_outer, long long max_inner, double* hist,
                                  int hist_cap, int* outer_steps, int* iterations,
                                  double* rel_residual, int* status) {
    return solve_multi_mixed_impl(h, method, nrhs, X, B, rel_tol, inner_rel_tol, max_outer,
                                  max_inner, hist, hist_cap, outer_steps, iterations,
                                  rel_residual, status, false);
}

GIS_API int gis_solve_multi_mixed_device(gis_handle h, int method, int nrhs, double* d_X,
                                         const double* d_B, double rel_tol,
                                         double inner_rel_tol, long long max_outer,
                                         long long max_inner, double* hist, int hist_cap,
                                         int* outer_steps, int* iterations,
                                         double* rel_residual, int* status) {
    return solve_multi_mixed_impl(h, method, nrhs, d_X, d_B, rel_tol, inner_rel_tol, max_outer,
                                  max_inner, hist, hist_cap, outer_steps, iterations,
                                  rel_residual, status, true);
}

static int adjoint_impl(gis_handle h, int method, int nrhs, const double* W, const double* x,
                        double* lambda, double* d_diag, double* d_lower, double* d_upper,
                        const gis_params* params, int* iterations, int* status, bool dev) {
    if (!h || !W || !x || !lambda) return GIS_ERR_ARG;
    if (nrhs < 1 || nrhs > 8) return fail(h, GIS_ERR_ARG, "nrhs must be in 1..8");
    if (method != GIS_CG && method != GIS_BICG)
        return fail(h, GIS_ERR_ARG, "the adjoint solve supports GIS_CG and GIS_BICG");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    const size_t N = (size_t)h->N, M = (size_t)h->M, k = (size_t)nrhs;
    // staging: lambda | W | x | d_diag | d_lower | d_upper
    const size_t tot = 2 * k * N + N + k * N + 2 * k * M;
    GIS_HIP(h, h->io2.ensure(sizeof(double) * tot));
    double* dl = h->io2.as<double>();
    double* dw = dl + k * N;
    double* dxf = dw + k * N;
    double* gd = dxf + N;
    double* glo = gd + k * N;
    double* gup = glo + k * M;
    const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GIS_HIP(h, hipMemcpyAsync(dw, W, sizeof(double) * k * N, kin, h->stream));
    GIS_HIP(h, hipMemcpyAsync(dxf, x, sizeof(double) * N, kin, h->stream));
    GIS_HIP(h, hipMemsetAsync(dl, 0, sizeof(double) * k * N, h->stream));
    SolveArgs a = to_args(method, params, 1);
    const int st = device_solve_multi(h, method, nrhs, true, a, dl, dw, iterations, nullptr,
                                      nullptr, status);
    if (st != GIS_OK && st < GIS_BREAKDOWN) return st;
    GIS_TRY(device_sddmm_fold(h, nrhs, dl, dxf, gd, glo, gup));
    GIS_HIP(h, hipMemcpyAsync(lambda, dl, sizeof(double) * k * N, kout, h->stream));
    if (d_diag) GIS_HIP(h, hipMemcpyAsync(d_diag, gd, sizeof(double) * k * N, kout, h->stream));
    if (d_lower && M)
        GIS_HIP(h, hipMemcpyAsync(d_lower, glo, sizeof(double) * k * M, kout, h->stream));
    if (d_upper && M)
        GIS_HIP(h, hipMemcpyAsync(d_upper, gup, sizeof(double) * k * M, kout, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    return st;
}

GIS_API int gis_adjoint(gis_handle h, int method, int nrhs, const double* W, const double* x,
                        double* lambda, double* d_diag, double* d_lower, double* d_upper,
                        const gis_params* params, int* iterations, int* status) {
    return adjoint_impl(h, method, nrhs, W, x, lambda, d_diag, d_lower, d_upper, params,
                        iterations, status, false);
}

GIS_API int gis_adjoint_device(gis_handle h, int method, int nrhs, const double* d_W,
                               const double* d_x, double* d_lambda, double* d_diag,
                               double* d_lower, double* d_upper, const gis_params* params,
                               int* iterations, int* status) {
    return adjoint_impl(h, method, nrhs, d_W, d_x, d_lambda, d_diag, d_lower, d_upper, params,
                        iterations, status, true);
}

GIS_API int gis_batch_set_pattern(gis_handle h, int n_systems, const int* n_cells,
                                  const int* n_faces, const int* lower_addr,
                                  const int* upper_addr) {
    if (!h || !n_cells || !n_faces) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    return batch_set_pattern(h, n_systems, n_cells, n_faces, lower_addr, upper_addr);
}

GIS_API int gis_batch_set_values(gis_handle h, const double* diag, const double* lower,
                                 const double* upper) {
    if (!h || !diag) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    return batch_set_values(h, diag, lower, upper, false);
}

GIS_API int gis_batch_set_values_device(gis_handle h, const double* d_diag,
                                        const double* d_lower, const double* d_upper) {
    if (!h || !d_diag) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    return batch_set_values(h, d_diag, d_lower, d_upper, true);
}

GIS_API int gis_batch_solve(gis_handle h, double* x, const double* b, const double* error,
                            long long max_steps, int use_graph, int* iterations,
                            double* final_residual, int* status) {
    if (!h || !x || !b || !error) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    return batch_solve(h, x, b, false, error, max_steps, use_graph, iterations, final_residual,
                       status);
}

GIS_API int gis_batch_solve_device(gis_handle h, double* d_x, const double* d_b,
                                   const double* error, long long max_steps, int use_graph,
                                   int* iterations, double* final_residual, int* status) {
    if (!h || !d_x || !d_b || !error) return GIS_ERR_ARG;
    DeviceGuard g(h->device);
    return batch_solve(h, d_x, d_b, true, error, max_steps, use_graph, iterations,
                       final_residual, status);
}

namespace {

// method: GIS_CG .. GIS_RICHARDSON, kMethodPcg0 + preconditioner for PCG, or
// kMethodPbicgDiag (the public entries check their own ranges)
int foam_solve_impl(gis_context* h, int method, double* psi, const double* source,
                    double tolerance, double rel_tol, long long min_iter, long long max_iter,
                    gis_performance* perf) {
    if (!h || !psi || !source) return GIS_ERR_ARG;
    if (!(rel_tol >= 0.0) || min_iter < 0)
        return fail(h, GIS_ERR_ARG, "rel_tol and min_iter must be non-negative");
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    DeviceGuard g(h->device);
    double nf = 0.0;
    GIS_TRY(norm_factor_impl(h, source, psi, &nf, false));
    gis_params p;
    gis_default_params(&p);
    p.error = tolerance * nf;   // cuFoamCG.C:68-74: error * norm_factor
    p.max_steps = max_iter;
    p.omega = 0.0;              // cuFoamRichardson.C:66
    p.power_steps = 30;
    p.int_bound = 1e-10;        // cuFoamBiCG.C:68
    SolveArgs a = to_args(method, &p, 1);
    a.rel_tol = rel_tol;
    a.min_steps = min_iter;
    int len = 0;
    double n0 = 0.0, nfin = 0.0;
    const int st = solve_staged(h, a, psi, source, false, nullptr, 0, &len, &n0, &nfin);
    if (st != GIS_OK && st < GIS_BREAKDOWN) return st;
    if (perf) {
        perf->initial_residual = n0;
        perf->final_residual = nfin;
        perf->n_iterations = len - 1;
        perf->status = st;
        perf->norm_factor = nf;
    }
    return st;
}

}  // namespace

GIS_API int gis_foam_solve(gis_handle h, int method, double* psi, const double* source,
                           double tolerance, long long max_iter, gis_performance* perf) {
    if (h && (method < GIS_CG || method > GIS_RICHARDSON))
        return fail(h, GIS_ERR_ARG, "unknown method");
    return foam_solve_impl(h, method, psi, source, tolerance, 0.0, 0, max_iter, perf);
}

GIS_API int gis_foam_solve_pcg(gis_handle h, int precond, double* psi, const double* source,
                               double tolerance, double rel_tol, long long min_iter,
                               long long max_iter, gis_performance* perf) {
    if (h && (precond < GIS_PRECOND_NONE || precond > GIS_PRECOND_DIC))
        return fail(h, GIS_ERR_ARG, "unknown preconditioner");
    return foam_solve_impl(h, kMethodPcg0 + precond, psi, source, tolerance, rel_tol, min_iter,
                           max_iter, perf);
}

GIS_API int gis_foam_solve_pbicg(gis_handle h, int precond, double* psi, const double* source,
                                 double tolerance, double rel_tol, long long min_iter,
                                 long long max_iter, gis_performance* perf) {
    int method = 0;
    if (h) GIS_TRY(pbicg_method(h, precond, &method));
    return foam_solve_impl(h, method, psi, source, tolerance, rel_tol, min_iter, max_iter, perf);
}

GIS_API int gis_foam_solve_controls(gis_handle h, int method, double* psi,
                                    const double* source, double tolerance, double rel_tol,
                                    long long min_iter, long long max_iter,
                                    gis_performance* perf) {
    if (h && (method < GIS_CG || method > GIS_RICHARDSON))
        return fail(h, GIS_ERR_ARG, "unknown method");
    return foam_solve_impl(h, method, psi, source, tolerance, rel_tol, min_iter, max_iter, perf);
}

}  // extern "C"
