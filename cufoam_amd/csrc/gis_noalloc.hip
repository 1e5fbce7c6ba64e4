// NormFactor without any device allocation (gis_norm_factor_device_noalloc).
//
// gis_norm_factor_device reduces through the handle's workspace: its level-0
// and level-1 slots, the global partials and the device KState, all grown on
// demand with hipMalloc, plus a staging copy of b and x.  Callers that may not
// allocate (code under stream capture, serving loops with a fixed memory
// budget, harnesses that forbid it) get the same kernels here with that
// workspace in one file-scope device array, `scratch`, that the code object
// brings with it: each GPU holds its own copy from module load to process end.
//
// Host code never passes `scratch` itself to a kernel or a copy (its host-side
// name is a stand-in, not a device address): the per-device address comes from
// hipGetSymbolAddress at set-up (noalloc_setup, from gis_create) and is kept in
// g_dev[device].base.
//
// Layout of the array: [KState (256 B) | reduction workspace for up to
// kNoAllocMaxRows rows (reduction_layout)].  The slots must hold the empty
// pattern before a launch; set-up fills them, every completed reduction leaves
// the slots it used empty again, and `clean` records the prefix of the slot
// region known to be empty (a call with a larger tree, or after a reduction
// that gave up, refills first).
#include <mutex>

#include "../../include/gis.h"
#include "gis_internal.hpp"

namespace gis {
namespace {

constexpr int kNoAllocMaxRows = 1 << 25;   // 33,554,432 rows (a 322^3 mesh)
constexpr int kMaxDevices = 64;

// doubles of the reduction workspace for `rows` rows (as reduction_layout)
constexpr size_t red_doubles(int rows) {
    long long count = (rows + kTile - 1) / kTile;
    long long c[kMaxLevels] = {};
    int L = 0;
    for (; L < kMaxLevels; ++L) {
        c[L] = count;
        if (count == 1) break;
        count = (count + kGroup - 1) / kGroup;
    }
    size_t n = (size_t)(c[0] + (L >= 1 ? c[1] : 1)) * kMaxNV;
    for (int j = 2; j <= L; ++j) n += (size_t)c[j] * kMaxNV;
    return n;
}

constexpr size_t kStateDoubles = 32;   // 256 B
static_assert(sizeof(KState) <= kStateDoubles * sizeof(double), "KState fits its 256 B");
constexpr size_t kRedDoubles = red_doubles(kNoAllocMaxRows);
constexpr size_t kScratchDoubles = kStateDoubles + kRedDoubles;   // about 4.3 MB

static __device__ __align__(256) double scratch[kScratchDoubles];

struct DevScratch {
    double* base = nullptr;   // device address of `scratch` on this device
    size_t clean = 0;         // leading doubles of the slot region holding the empty pattern
};

DevScratch g_dev[kMaxDevices];
std::mutex g_mu[kMaxDevices];

}  // namespace

// Once per device, with the device current (gis_create holds it): the array's
// address on this device, and its slot region filled with the empty pattern.
int noalloc_setup(gis_context* h) {
    const int d = h->device;
    if (d < 0 || d >= kMaxDevices) return kOK;   // the entry point rejects such devices
    std::lock_guard<std::mutex> lk(g_mu[d]);
    DevScratch& s = g_dev[d];
    if (s.base) return kOK;
    void* p = nullptr;
    GIS_HIP(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    double* base = static_cast<double*>(p);
    GIS_TRY(fill_slots(h, base + kStateDoubles, kRedDoubles));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    s.base = base;
    s.clean = kRedDoubles;
    return kOK;
}

}  // namespace gis

using namespace gis;

extern "C" {

GIS_API int gis_norm_factor_device_noalloc(gis_handle h, const double* d_b, const double* d_x,
                                           double* nf) {
    if (!h || !d_b || !d_x || !nf) return GIS_ERR_ARG;
    if (!h->has_values) return fail(h, GIS_ERR_STATE, "values not set");
    if (h->N > kNoAllocMaxRows)
        return fail(h, GIS_ERR_SHAPE,
                    "norm_factor_noalloc: at most 33554432 rows (the scratch array's size)");
    const int d = h->device;
    if (d < 0 || d >= kMaxDevices) return fail(h, GIS_ERR_ARG, "norm_factor_noalloc: device id");
    DeviceGuard g(d);
    std::lock_guard<std::mutex> lk(g_mu[d]);
    DevScratch& s = g_dev[d];
    if (!s.base) return fail(h, GIS_ERR_STATE, "norm_factor_noalloc: scratch not set up");
    KState* st = reinterpret_cast<KState*>(s.base);
    double* red_base = s.base + kStateDoubles;
    RedWs red{};
    size_t ndbl = 0, nslot = 0;
    if (!reduction_layout(h->N, red_base, &red, &ndbl, &nslot) || ndbl > kRedDoubles)
        return fail(h, GIS_ERR_SHAPE, "norm_factor_noalloc: reduction workspace too large");
    const size_t need = nslot;
    if (s.clean < need) GIS_TRY(fill_slots(h, red_base, need));
    s.clean = 0;   // until this call's reductions are known to have completed
    GIS_TRY(launch_norm_factor(h, red, st, d_b, d_x));
    KState out;
    GIS_HIP(h, hipMemcpyAsync(&out, st, sizeof(KState), hipMemcpyDeviceToHost, h->stream));
    GIS_HIP(h, hipStreamSynchronize(h->stream));
    if (out.status == kErrSync)
        return fail(h, GIS_ERR_SYNC, "norm_factor_noalloc: reduction did not complete");
    // the collectors left every slot of this layout empty; the partials above
    // it (levels 2 and up) now hold values
    s.clean = need;
    *nf = out.aux[1];
    return GIS_OK;
}

}  // extern "C"
