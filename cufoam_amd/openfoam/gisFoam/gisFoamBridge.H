/*
 * gisFoamBridge.H -- the solve sequence of the cuFoam<Name> wrappers over the
 * gis C ABI (include/gis.h), in plain C++ with no OpenFOAM type, so that it
 * builds and is tested without an OpenFOAM installation.
 *
 * The reference wrappers (cuFoamCG/cuFoamCG.C:38-89 and siblings) rebuild a
 * host CSR, allocate and upload it and download psi on every solve.  Here one
 * gis handle per mesh addressing keeps the device pattern; a solve uploads the
 * coefficients (gis_set_values: one gather on the device) and runs
 * gis_foam_solve, which applies NormFactor, error = tolerance * nf, the
 * per-wrapper constants and the solverPerformance fields of the reference.
 */
#ifndef GIS_FOAM_BRIDGE_H
#define GIS_FOAM_BRIDGE_H

#include <cstddef>
#include <map>
#include <mutex>
#include <string>

#include "gis.h"

namespace gisFoam {

// lduAddressing as plain arrays (OpenFOAM labels are 32-bit: WM_LABEL_SIZE=32)
struct Addressing {
    const void* owner;      // identity of the addressing object (the handle cache key)
    int nCells;
    int nFaces;
    const int* lowerAddr;   // owner cell of each face, faces ordered by owner
    const int* upperAddr;   // neighbour cell of each face
};

struct Coefficients {
    const double* diag;     // nCells
    const double* lower;    // nFaces: A(neighbour, owner)
    const double* upper;    // nFaces: A(owner, neighbour)
};

struct Controls {
    double tolerance = 0.0;   // absolute, on the NormFactor-scaled L2 residual
    long long maxIter = 0;    // <= 0: unbounded
    double relTol = 0.0;      // OpenFOAM's relTol, used when `openfoam` is set
    long long minIter = 0;    // OpenFOAM's minIter, used when `openfoam` is set
    bool openfoam = false;    // false: tolerance and maxIter only, as the reference wrappers
    int device = 0;
};

struct Outcome {
    gis_performance perf{};   // initial/final residual, iterations, status, NormFactor
    int status = GIS_OK;
    bool fatal = false;       // an error: psi is not a solution (FatalError)
    bool warn = false;        // BiCG breakdown or a NaN residual: psi is valid but suspect
    std::string message;
};

// GIS_OK and the solve outcomes (breakdown, iteration cap, NaN) are results;
// everything else, including GIS_ERR_SYNC, is an error.
inline bool isError(int st) { return st != GIS_OK && (st < GIS_BREAKDOWN || st == GIS_ERR_SYNC); }

// One handle per mesh addressing.  A cached handle is reused only while the
// addressing's sizes and array pointers are unchanged, so an addressing object
// freed and another one allocated at the same address (a topology change)
// gets its pattern rebuilt.  Handles live until clear() or process exit (the
// GPU memory goes with the process; freeing it from static destructors could
// run after the HIP runtime has shut down).
class Handles {
public:
    static Handles& global() {
        static Handles* h = new Handles();
        return *h;
    }

    int get(const Addressing& a, int device, gis_handle* out, std::string* msg) {
        std::lock_guard<std::mutex> lock(mu_);
        Entry& e = map_[a.owner];
        if (!e.h || e.device != device) {
            if (e.h) gis_destroy(e.h);
            e = Entry();
            const int st = gis_create(&e.h, device);
            if (st != GIS_OK) {
                e.h = nullptr;
                *msg = std::string("gis_create: ") + gis_status_string(st);
                return st;
            }
            e.device = device;
        }
        if (e.nCells != a.nCells || e.nFaces != a.nFaces || e.lower != a.lowerAddr ||
            e.upper != a.upperAddr) {
            e.nCells = -1;
            const int st = gis_set_pattern(e.h, a.nCells, a.nFaces, a.lowerAddr, a.upperAddr);
            if (st != GIS_OK) {
                *msg = std::string("gis_set_pattern: ") + gis_last_error(e.h);
                return st;
            }
            e.nCells = a.nCells;
            e.nFaces = a.nFaces;
            e.lower = a.lowerAddr;
            e.upper = a.upperAddr;
            ++patterns_;
        }
        *out = e.h;
        return GIS_OK;
    }

    void clear() {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto& kv : map_)
            if (kv.second.h) gis_destroy(kv.second.h);
        map_.clear();
    }

    std::size_t size() {
        std::lock_guard<std::mutex> lock(mu_);
        return map_.size();
    }

    // gis_set_pattern calls so far (device pattern builds)
    long long patternBuilds() {
        std::lock_guard<std::mutex> lock(mu_);
        return patterns_;
    }

private:
    struct Entry {
        gis_handle h = nullptr;
        int device = -1;
        int nCells = -1;
        int nFaces = -1;
        const int* lower = nullptr;
        const int* upper = nullptr;
    };
    std::map<const void*, Entry> map_;
    std::mutex mu_;
    long long patterns_ = 0;
};

// Which gis entry a wrapper's solve ends in.
enum class Family {
    Plain,   // gis_foam_solve / gis_foam_solve_controls with a GIS_<method> id
    PCG,     // gis_foam_solve_pcg with a GIS_PRECOND_<name> id
    PBiCG,   // gis_foam_solve_pbicg with GIS_PRECOND_NONE, _DIAGONAL or _DILU
};

// cuFoam<Name>::solve: psi (host, nCells) is updated in place.  The handle of
// the addressing, gis_set_values, then the family's entry; relTol / minIter
// are applied when `openfoam` is set.
inline Outcome solveAs(Family family, int id, const Addressing& a, const Coefficients& c,
                       double* psi, const double* source, const Controls& ctl) {
    Outcome o;
    gis_handle h = nullptr;
    int st = Handles::global().get(a, ctl.device, &h, &o.message);
    if (st == GIS_OK) {
        st = gis_set_values(h, c.diag, c.lower, c.upper);
        if (st != GIS_OK) o.message = std::string("gis_set_values: ") + gis_last_error(h);
    }
    if (st == GIS_OK) {
        const double relTol = ctl.openfoam ? ctl.relTol : 0.0;
        const long long minIter = ctl.openfoam ? ctl.minIter : 0;
        const char* entry = "gis_foam_solve";
        if (family == Family::PCG) {
            entry = "gis_foam_solve_pcg";
            st = gis_foam_solve_pcg(h, id, psi, source, ctl.tolerance, relTol, minIter,
                                    ctl.maxIter, &o.perf);
        } else if (family == Family::PBiCG) {
            entry = "gis_foam_solve_pbicg";
            st = gis_foam_solve_pbicg(h, id, psi, source, ctl.tolerance, relTol, minIter,
                                      ctl.maxIter, &o.perf);
        } else if (ctl.openfoam) {
            entry = "gis_foam_solve_controls";
            st = gis_foam_solve_controls(h, id, psi, source, ctl.tolerance, relTol, minIter,
                                         ctl.maxIter, &o.perf);
        } else {
            st = gis_foam_solve(h, id, psi, source, ctl.tolerance, ctl.maxIter, &o.perf);
        }
        if (isError(st)) o.message = std::string(entry) + ": " + gis_last_error(h);
    }
    o.status = st;
    o.fatal = isError(st);
    o.warn = (st == GIS_BREAKDOWN || st == GIS_NAN);
    if (o.warn) o.message = gis_status_string(st);
    return o;
}

inline Outcome solve(int method, const Addressing& a, const Coefficients& c, double* psi,
                     const double* source, const Controls& ctl) {
    return solveAs(Family::Plain, method, a, c, psi, source, ctl);
}

// Preconditioned CG (GIS_PRECOND_NONE / DIAGONAL / DIC / GAMG / GAMG32).
inline Outcome solvePcg(int precond, const Addressing& a, const Coefficients& c, double* psi,
                        const double* source, const Controls& ctl) {
    return solveAs(Family::PCG, precond, a, c, psi, source, ctl);
}

// Preconditioned BiCG (GIS_PRECOND_NONE / DIAGONAL / DILU).
inline Outcome solvePbicg(int precond, const Addressing& a, const Coefficients& c,
                          double* psi, const double* source, const Controls& ctl) {
    return solveAs(Family::PBiCG, precond, a, c, psi, source, ctl);
}

}  // namespace gisFoam

#endif  // GIS_FOAM_BRIDGE_H
