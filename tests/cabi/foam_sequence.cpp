// A compiled C++ caller of the gis C ABI: includes only include/gis.h and the
// wrappers' OpenFOAM-free bridge (cufoam_amd/openfoam/gisFoam/gisFoamBridge.H),
// links libgis_amd.so, and runs the cuFoam<Name>::solve sequence of the
// wrappers (handle per mesh addressing, gis_set_pattern once, gis_set_values
// and gis_foam_solve per solve; cuFoamCG.C:38-89 in the reference).
//
//   foam_sequence layout            struct sizes and offsets as JSON (no GPU call)
//   foam_sequence solve IN OUT      systems from IN, results to OUT (below)
//
// IN:  int32 K, then K times: int32 method, N, M, maxIter; float64 tolerance;
//      (method < 100: a GIS_<method> id; 100 + p: PCG, 200 + p: PBiCG with
//      the GIS_PRECOND_ id p)
//      int32 lowerAddr[M], upperAddr[M]; float64 diag[N], lower[M], upper[M],
//      source[N], psi[N].
// OUT: int64 pattern builds, then per system: int32 status, nIterations,
//      repeat_identical; float64 initial, final, normFactor; float64 psi[N].
// Every system is solved twice from its input psi under one handle-cache key
// (a new mesh behind the same addressing object): the first solve of a system
// rebuilds the pattern, the second reuses it and must give the same bits.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gis.h"
#include "gisFoamBridge.H"

namespace {

template <typename T>
bool get(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

template <typename T>
void put(FILE* f, const T* p, size_t n) { fwrite(p, sizeof(T), n, f); }

int layout() {
    std::printf("{\"gis_params\": [%zu, %zu, %zu, %zu, %zu, %zu, %zu],\n", sizeof(gis_params),
                offsetof(gis_params, error), offsetof(gis_params, max_steps),
                offsetof(gis_params, omega), offsetof(gis_params, power_steps),
                offsetof(gis_params, int_bound), offsetof(gis_params, detail_type));
    std::printf(" \"gis_performance\": [%zu, %zu, %zu, %zu, %zu, %zu],\n",
                sizeof(gis_performance), offsetof(gis_performance, initial_residual),
                offsetof(gis_performance, final_residual),
                offsetof(gis_performance, n_iterations), offsetof(gis_performance, status),
                offsetof(gis_performance, norm_factor));
    std::printf(" \"gis_schedule\": [%zu, %zu, %zu, %zu, %zu, %zu, %zu, %zu]}\n",
                sizeof(gis_schedule), offsetof(gis_schedule, spmv64),
                offsetof(gis_schedule, spmv32), offsetof(gis_schedule, block_iters),
                offsetof(gis_schedule, fp64_direction), offsetof(gis_schedule, fp32_direction),
                offsetof(gis_schedule, coarsen64), offsetof(gis_schedule, coarsen32));
    return 0;
}

int run(const char* in_path, const char* out_path) {
    FILE* in = std::fopen(in_path, "rb");
    if (!in) return 2;
    int32_t K = 0;
    if (!get(in, &K, 1)) return 2;
    static int addressing_object;   // one cache key for every system
    std::vector<std::vector<char>> keep;
    struct Result {
        int32_t status, iters, same;
        double init, fin, nf;
        std::vector<double> psi;
    };
    std::vector<Result> results;
    for (int s = 0; s < K; ++s) {
        int32_t hdr[4];
        double tol;
        if (!get(in, hdr, 4) || !get(in, &tol, 1)) return 2;
        const int method = hdr[0], N = hdr[1], M = hdr[2], max_iter = hdr[3];
        std::vector<int32_t> la(M), ua(M);
        std::vector<double> diag(N), lower(M), upper(M), source(N), psi0(N);
        if (!get(in, la.data(), M) || !get(in, ua.data(), M) || !get(in, diag.data(), N) ||
            !get(in, lower.data(), M) || !get(in, upper.data(), M) ||
            !get(in, source.data(), N) || !get(in, psi0.data(), N))
            return 2;
        const gisFoam::Addressing a{&addressing_object, N, M, la.data(), ua.data()};
        const gisFoam::Coefficients c{diag.data(), lower.data(), upper.data()};
        gisFoam::Controls ctl;
        ctl.tolerance = tol;
        ctl.maxIter = max_iter;
        Result r;
        std::vector<double> psi = psi0;
        const gisFoam::Family fam = method >= 200   ? gisFoam::Family::PBiCG
                                    : method >= 100 ? gisFoam::Family::PCG
                                                    : gisFoam::Family::Plain;
        const int id = method % 100;
        gisFoam::Outcome o1 = gisFoam::solveAs(fam, id, a, c, psi.data(), source.data(), ctl);
        std::vector<double> psi2 = psi0;
        gisFoam::Outcome o2 = gisFoam::solveAs(fam, id, a, c, psi2.data(), source.data(), ctl);
        if (o1.fatal || o2.fatal) {
            std::fprintf(stderr, "system %d: %s%s\n", s, o1.message.c_str(), o2.message.c_str());
            return 3;
        }
        r.same = (std::memcmp(psi.data(), psi2.data(), sizeof(double) * N) == 0 &&
                  o1.perf.final_residual == o2.perf.final_residual &&
                  o1.perf.n_iterations == o2.perf.n_iterations) ? 1 : 0;
        r.status = o2.status;
        r.iters = o2.perf.n_iterations;
        r.init = o2.perf.initial_residual;
        r.fin = o2.perf.final_residual;
        r.nf = o2.perf.norm_factor;
        r.psi = psi2;
        results.push_back(r);
    }
    std::fclose(in);
    FILE* out = std::fopen(out_path, "wb");
    if (!out) return 2;
    const long long builds = gisFoam::Handles::global().patternBuilds();
    put(out, &builds, 1);
    for (const Result& r : results) {
        put(out, &r.status, 1);
        put(out, &r.iters, 1);
        put(out, &r.same, 1);
        put(out, &r.init, 1);
        put(out, &r.fin, 1);
        put(out, &r.nf, 1);
        put(out, r.psi.data(), r.psi.size());
    }
    std::fclose(out);
    gisFoam::Handles::global().clear();
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "layout") == 0) return layout();
    if (argc >= 4 && std::strcmp(argv[1], "solve") == 0) return run(argv[2], argv[3]);
    std::fprintf(stderr, "usage: %s layout | solve IN OUT\n", argv[0]);
    return 1;
}
