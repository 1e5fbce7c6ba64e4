#!/usr/bin/env python3
"""Benchmark: CG time to 1e-10 relative residual on the 256^3 Poisson system,
FP64 vs mixed precision, and CSR SpMV HBM bandwidth, on one MI355X.

A step is one complete solve from x0 = 0 of the 256^3 7-point FV Laplacian
(shift 1e-4, seed 0, b ~ U[-1,1]) until ||b - A x||_2 / ||b||_2 <= 1e-10, with
matrix and vectors already resident in HBM.  `value` is the mixed-precision
solve time (FP32 Krylov loop, FP64 residual/correction/norms); the plain FP64
CG time of the same solve is reported beside it.  The roofline entry is for
the dominant kernel of the mixed solve (the FP32 SpMV + (p, q) kernel of the
split schedule, or the fused SpMV + direction + correction kernel), timed live
with HIP events on its launch stream.  `achieved` and `frac` count the bytes the
kernel must move in the layout it runs on (gis_get_layout: the stored matrix of
the SellD slices, x read once, q written once); `achieved_csr` / `frac_csr` count
the CSR bytes of the same SpMV (an int32 column index per nonzero), which SellD
does not store; `traffic` is the HBM bytes the kernel was measured to move.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
# the dominant kernel of a timed solve is timed live with HIP events on one
# launch in TIMING_PERIOD: an event pair adds a few microseconds of launch gap,
# which would otherwise land in the measured solve time.  Odd, so that the
# samples alternate between the two kinds of k_mx1 launch (with and without the
# paired correction update, about 168 and 139 us at 256^3)
TIMING_PERIOD = 15
sys.path.insert(0, ROOT)

METRIC = "CG time-to-1e-10 rel resid, 256^3 Poisson: FP64 vs mixed-prec; SpMV HBM GB/s"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(N, nnz, sv, sx):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row offsets + x read + y write."""
    return nnz * (sv + 4) + 4 * (N + 1) + 2 * N * sx


def mx1_bytes(N, nnz, direction):
    """CSR algorithmic bytes of the FP32 SpMV-carrying iteration kernel, per launch on
    average.  Fused form (k_mx1): FP32 matrix (values + columns + row offsets) and
    FP32 vector passes: r and p_{k-1} read, p_k and q written, and every other launch
    (the deferred correction goes into xc in pairs) xc and p_{k-2} read and xc
    written, 5.5 passes on average.  Split form (k_mx1s): the matrix, p_k read once
    and q written, i.e. the bytes of one FP32 SpMV."""
    if direction:
        return spmv_bytes(N, nnz, 4, 4)
    return nnz * (4 + 4) + 4 * (N + 1) + 22 * N


def mx1_layout_bytes(g, N, schedule, direction):
    """The same kernel's bytes in the layout it runs on: one FP32 SpMV of the stored
    matrix (Gis.spmv_bytes) plus, for the fused form, the extra FP32 vector passes:
    r and p_{k-1} read, and every other launch xc and p_{k-2} read and xc written
    (3.5 passes on average)."""
    b = g.spmv_bytes(4, 4, schedule)
    return b if direction else b + 14 * N


def cpu_baseline(sys_, A, x_gpu_fp64, x_gpu_mixed, iters_fp64: int, sample_iters: int):
    """FP64 CG of the same system on the host, run in full (not extrapolated): the
    oracle's restatement of SolverSingleGPU.h:366-445 on all OpenMP threads, with the
    reduction tree of the GPU path, so its iterates are the GPU FP64 CG's bit for bit.
    Its solution is the FP64 oracle solve the mixed result is measured against.
    OpenFOAM's PCG (preconditioner none) restated on all threads runs the same number
    of iterations beside it; one-thread figures are extrapolated from short samples."""
    import numpy as np
    from oracle import oracle as O
    threads = O.par_threads()
    bnorm = float(np.linalg.norm(sys_.b))
    t0 = time.perf_counter()
    ref = O.cg_par(A, sys_.b, sys_.x0, 1e-10 * bnorm, 0, hist_cap=iters_fp64 + 2)
    t_full = time.perf_counter() - t0
    xo = ref.x
    err_m = float(np.linalg.norm(x_gpu_mixed - xo) / np.linalg.norm(xo))
    _, it_of, res_of, t_of = O.pcg_openfoam_omp(A, sys_.b, 0.0, iters_fp64)
    t_ldu, _ = O.time_cg_ldu(sys_, sample_iters)
    _, _, _, t_of1 = O.pcg_openfoam(sys_, 0.0, sample_iters)
    return {
        "value": round(t_full, 3),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole FP64 CG solve of the same 256^3 system to ||r|| <= 1e-10 ||b|| "
                  f"({ref.iterations} iterations) on {threads} OpenMP threads: the oracle's "
                  f"restatement with the GPU path's reduction tree",
        "iterations": ref.iterations,
        "gpu_fp64_bit_exact": bool(np.array_equal(x_gpu_fp64, xo)),
        "rel_error_mixed_vs_oracle_fp64": err_m,
        "rel_residual_oracle": O.residual_norm_par(A, sys_.b, xo) / bnorm,
        "openfoam_pcg_none_all_threads": {
            "s": round(t_of, 3), "iterations": it_of, "scaled_l1_residual": res_of,
            "note": "OpenFOAM's PCG loop restated, row-parallel matvec, same iteration count"},
        "one_thread_extrapolated": {
            "openfoam_pcg_none_s": round(t_of1 / sample_iters * iters_fp64, 2),
            "ldu_face_loop_cg_s": round(t_ldu / sample_iters * iters_fp64, 2),
            "sample": f"{sample_iters} iterations each, scaled to {iters_fp64}"},
    }


def batch_throughput(reps: int, use_graph: bool = True):
    """64 SPD systems with edges n_s ~ U{16..64} (seed 0), each solved by CG from x0 = 0
    to ||r||_2 <= 1e-8 ||b||_2: one batched solve of all 64 against the same 64 solves
    done one after another through single-system handles (values resident in both)."""
    import numpy as np
    import torch
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    edges = mesh.batch_edges(64, 16, 64, seed=0)
    systems = [mesh.laplacian(int(e), seed=100 + i, shift=1e-3) for i, e in enumerate(edges)]
    err = np.array([1e-8 * np.linalg.norm(s.b) for s in systems])
    g = Gis(0)
    g.batch_set_systems(systems)
    b = torch.from_numpy(np.concatenate([s.b for s in systems])).cuda()
    singles = []
    for s in systems:
        h = Gis(0)
        h.set_system(s)
        singles.append((h, torch.from_numpy(s.b).cuda()))

    def run_batch():
        return g.batch_solve(b, torch.zeros_like(b), err, 100000, use_graph=use_graph)

    def run_seq():
        its = []
        for k, (h, bs) in enumerate(singles):
            _, hist, _ = h.solve("CG", bs, torch.zeros_like(bs), error=err[k],
                                 max_steps=100000, hist_cap=1)
            its.append(h.last_iterations)
        return its

    run_batch()
    run_seq()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, it_b, _, st_b = run_batch()
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        it_s = run_seq()
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) / reps
    out = {"systems": 64, "cells": int(sum(s.n_cells for s in systems)),
           "nnz": int(sum(s.nnz for s in systems)), "graph": use_graph,
           "batched_s": round(t_b, 5), "systems_per_s": round(64 / t_b, 2),
           "sequential_s": round(t_s, 5), "sequential_systems_per_s": round(64 / t_s, 2),
           "speedup": round(t_s / t_b, 3), "max_iterations": int(max(it_b)),
           "iterations_equal": bool(list(map(int, it_b)) == list(map(int, it_s))),
           "all_converged": bool(np.all(st_b == 0))}
    for h, _ in singles:
        h.close()
    g.close()
    return out


def multi_rhs_throughput(reps: int):
    """128^3 nonsymmetric convection-diffusion system with 3 right-hand sides (seed 0): one
    SpMM-BiCG solve of all three columns against three single-RHS BiCG solves, each column
    to ||r||_2 <= 1e-8 ||b_0||_2 from x0 = 0, everything resident in HBM; then FP64
    SpMM-BiCG against FP32/FP64 refinement (FP32 SpMM-BiCG inner solves, inner relTol 1e-4),
    both to relative residual 1e-10."""
    import numpy as np
    import torch
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    s = mesh.config_system("convdiff128x3")
    g = Gis(0)
    g.set_pattern(s.n_cells, s.lower_addr, s.upper_addr)
    g.set_values(s.diag, s.lower, s.upper)
    B = torch.from_numpy(np.ascontiguousarray(s.b)).cuda()
    err = 1e-8 * float(np.linalg.norm(s.b[0]))

    def multi():
        return g.solve_multi("BiCG", B, torch.zeros_like(B), error=err, max_steps=20000)

    def single():
        its = []
        for k in range(B.shape[0]):
            g.solve("BiCG", B[k], torch.zeros_like(B[k]), error=err, max_steps=20000, hist_cap=1)
            its.append(g.last_iterations)
        return its

    multi()
    single()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, it_m, _, st_m = multi()
    torch.cuda.synchronize()
    t_m = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        it_s = single()
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) / reps
    # to 1e-10 relative: FP64 SpMM-BiCG against FP32/FP64 refinement (SpMM-BiCG inner solves)
    bn = np.linalg.norm(s.b, axis=1)
    err10 = 1e-10 * float(bn.min())
    g.solve_multi("BiCG", B, torch.zeros_like(B), error=err10, max_steps=20000)
    g.solve_multi_mixed("BiCG", B, torch.zeros_like(B), rel_tol=1e-10, inner_rel_tol=1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X64, it64, _, _ = g.solve_multi("BiCG", B, torch.zeros_like(B), error=err10, max_steps=20000)
    torch.cuda.synchronize()
    t_64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    Xm, hm, itm, relm, stm = g.solve_multi_mixed("BiCG", B, torch.zeros_like(B), rel_tol=1e-10,
                                                 inner_rel_tol=1e-4)
    torch.cuda.synchronize()
    t_mx = time.perf_counter() - t0
    R = torch.empty_like(X64[0])
    rel64 = []
    for k in range(B.shape[0]):
        g.spmv(X64[k], R)
        rel64.append(float(torch.linalg.vector_norm(B[k] - R)) / float(bn[k]))
    g.close()
    return {"n_cells": s.n_cells, "rhs": 3, "spmm_bicg_s": round(t_m, 5),
            "three_single_bicg_s": round(t_s, 5), "speedup": round(t_s / t_m, 3),
            "iterations": [int(v) for v in it_m],
            "iterations_equal_single": [int(v) for v in it_m] == [int(v) for v in it_s],
            "all_converged": bool(np.all(st_m == 0)),
            "to_1e-10": {"fp64_spmm_bicg_s": round(t_64, 5),
                         "fp64_iterations": [int(v) for v in it64],
                         "fp64_rel_residual": rel64,
                         "mixed_refinement_s": round(t_mx, 5),
                         "mixed_inner_iterations": [int(v) for v in itm],
                         "mixed_fp64_steps": [len(h) for h in hm],
                         "mixed_rel_residual": [float(v) for v in relm],
                         "mixed_status": [int(v) for v in stm],
                         "speedup_mixed": round(t_64 / t_mx, 3)}}


def dropin(cpu_sample: int):
    """The OpenFOAM drop-in path at 128^3 (cuFoam<Name>::solve through gis_foam_solve:
    host psi/source in and out, error = tolerance * NormFactor, tolerance 1e-8, x0 = 0):
    CG, Jacobi and Richardson on the SPD system (shift 1e-3), BiCG and MRJacobi on the
    nonsymmetric one (shift 0.1).  The handle runs the schedule gis_set_pattern picks
    for the pattern (no tuning call, as in the wrappers).  Each solve is timed with and
    without the per-solve coefficient upload (gis_set_values).  CPU: OpenFOAM's PCG
    (for cuFoamCG) and PBiCG (for cuFoamBiCG) with preconditioner none, restated on all
    host threads and run in full to the same tolerance; for the other wrappers, which
    have no OpenFOAM counterpart, the oracle's same method on one host thread, timed on
    a bounded sample of iterations and extrapolated to the GPU's count."""
    import numpy as np
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    systems = {"poisson128": mesh.config_system("poisson128"),
               "convdiff128": mesh.config_system("convdiff128")}
    cases = [("CG", "poisson128", 5000), ("Jacobi", "poisson128", 2000),
             ("Richardson", "poisson128", 2000), ("BiCG", "convdiff128", 5000),
             ("MRJacobi", "convdiff128", 5000)]
    out = {}
    try:
        from oracle import oracle as O
    except Exception:
        O = None
    for name, sysname, max_iter in cases:
        s = systems[sysname]
        g = Gis(0)
        g.set_pattern(s.n_cells, s.lower_addr, s.upper_addr)
        g.set_values(s.diag, s.lower, s.upper)
        sched = g.get_schedule()                         # the pattern's own schedule
        psi = np.zeros(s.n_cells)
        g.foam_solve(name, psi, s.b, 1e-8, max_iter)     # warm-up (code load, workspaces)
        psi = np.zeros(s.n_cells)
        t0 = time.perf_counter()
        g.set_values(s.diag, s.lower, s.upper)
        t1 = time.perf_counter()
        perf = g.foam_solve(name, psi, s.b, 1e-8, max_iter)
        t2 = time.perf_counter()
        g.close()
        rec = {"system": sysname, "max_iter": max_iter, "iterations": perf.n_iterations,
               "schedule": sched,
               "initial_residual": perf.initial_residual,
               "final_residual": perf.final_residual, "status": perf.status,
               "converged": bool(perf.final_residual <= 1e-8 * perf.norm_factor),
               "solve_s": round(t2 - t1, 5), "with_set_values_s": round(t2 - t0, 5),
               "us_per_iteration": round((t2 - t1) / max(perf.n_iterations, 1) * 1e6, 2)}
        if O is not None and cpu_sample > 0 and name in ("CG", "BiCG"):
            A = O.ldu2csr(s)
            f = O.pcg_openfoam_omp if name == "CG" else O.pbicg_openfoam_omp
            _, it_of, res_of, t_of = f(A, s.b, 1e-8, max_iter)
            rec["cpu_openfoam_all_threads"] = {
                "solver": "PCG" if name == "CG" else "PBiCG", "preconditioner": "none",
                "threads": O.baseline_threads(), "s": round(t_of, 4), "iterations": it_of,
                "scaled_l1_residual": res_of,
                "note": "OpenFOAM's loop restated, full run to tolerance 1e-8 on its own "
                        "residual (sum|r| / normFactor)"}
        elif O is not None and cpu_sample > 0:
            A = O.ldu2csr(s)
            t = []
            for k in (cpu_sample, 2 * cpu_sample):   # the difference drops set-up work
                t0 = time.perf_counter()
                O.solve(name, A, s.b, s.x0, 0.0, k, hist_cap=k + 1)
                t.append(time.perf_counter() - t0)
            per = (t[1] - t[0]) / cpu_sample
            rec["cpu_1thread_s"] = round(t[0] - cpu_sample * per + per * perf.n_iterations, 3)
            rec["cpu_sample"] = (f"the oracle's {name} for {cpu_sample} and {2 * cpu_sample} "
                                 f"iterations on one thread: set-up plus per-iteration time, "
                                 f"extrapolated to the GPU's iteration count")
        out[f"cuFoam{name}"] = rec
        log(f"[bench] drop-in cuFoam{name}: {rec}")
    return out


def pcg_128():
    """Preconditioners on the 128^3 SPD system (shift 1e-3, x0 = 0, to ||r|| <= 1e-10 ||b||):
    none, diagonal, DIC (OpenFOAM's incomplete Cholesky, sweeps level by level) and
    GAMG.  Each is timed once after a one-iteration call that builds its set-up."""
    import numpy as np
    import torch
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    s = mesh.config_system("poisson128")
    g = Gis(0)
    g.set_system(s)
    b = torch.from_numpy(s.b).cuda()
    err = 1e-10 * float(np.linalg.norm(s.b))
    out = {}
    for pc in ("none", "diagonal", "DIC", "GAMG"):
        t0 = time.perf_counter()
        g.solve_pcg(pc, b, torch.zeros_like(b), error=0.0, max_steps=1)
        t_setup = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist, st = g.solve_pcg(pc, b, torch.zeros_like(b), error=err, max_steps=20000,
                                  hist_cap=20001)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        it = len(hist) - 1
        out[pc] = {"s": round(t, 4), "iterations": it,
                   "ms_per_iteration": round(t / max(it, 1) * 1e3, 3),
                   "first_call_s": round(t_setup, 3), "status": st}
    g.close()
    log(f"[bench] pcg 128^3: {out}")
    return out


def pbicg_128():
    """Preconditioners for BiCG on the 128^3 nonsymmetric convection-diffusion system
    (x0 = 0, to ||r|| <= 1e-10 ||b||): none, diagonal and DILU (OpenFOAM's incomplete
    LU, sweeps level by level).  Each is timed once after a one-iteration call that
    builds its set-up."""
    import numpy as np
    import torch
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    s = mesh.config_system("convdiff128")
    g = Gis(0)
    g.set_system(s)
    b = torch.from_numpy(s.b).cuda()
    err = 1e-10 * float(np.linalg.norm(s.b))
    out = {}
    for pc in ("none", "diagonal", "DILU"):
        t0 = time.perf_counter()
        g.solve_pbicg(pc, b, torch.zeros_like(b), error=0.0, max_steps=1)
        t_setup = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist, st = g.solve_pbicg(pc, b, torch.zeros_like(b), error=err, max_steps=20000,
                                    hist_cap=20001)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        it = len(hist) - 1
        out[pc] = {"s": round(t, 4), "iterations": it,
                   "ms_per_iteration": round(t / max(it, 1) * 1e3, 3),
                   "first_call_s": round(t_setup, 3), "status": st}
    g.close()
    log(f"[bench] pbicg 128^3: {out}")
    return out


def unstructured_spmv():
    """SpMV where the offset tables do not apply: the 128^3 SPD system with its cells
    renumbered at random (every row's neighbours scattered over the whole vector)
    and then by reverse Cuthill-McKee (banded, as mesh tools order cells).  The
    schedule is the one gis_set_pattern picks; the structured numbering is timed
    beside them.  Bytes are those of the layout that runs (Gis.spmv_bytes)."""
    import numpy as np
    import torch
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    s = mesh.config_system("poisson128")
    N = s.n_cells
    rng = np.random.default_rng(7)
    rand = mesh.renumber(s, rng.permutation(N))
    adj = csr_matrix((np.ones(rand.n_faces), (rand.lower_addr, rand.upper_addr)), shape=(N, N))
    order = reverse_cuthill_mckee(adj + adj.T, symmetric_mode=True)
    perm = np.empty(N, np.int64)
    perm[order] = np.arange(N)
    rcm = mesh.renumber(rand, perm)
    out = {}
    for name, sys_ in (("structured", s), ("random", rand), ("rcm", rcm)):
        g = Gis(0)
        g.set_system(sys_)
        sched = g.get_schedule()
        lay = g.layout()
        rec = {"spmv_schedule": sched["spmv64"],
               "tabled_slot_share": round(1.0 - lay["untabled_slots"] / max(lay["sell_slots"], 1)
                                          if lay["table_slices"] else 0.0, 4),
               "sell_slots_per_nnz": round(lay["sell_slots"] / lay["nnz"], 4)}
        x64 = torch.randn(N, dtype=torch.float64, device="cuda")
        for tag, xx, sv in (("fp64", x64, 8), ("fp32", x64.float(), 4)):
            yy = torch.empty_like(xx)
            for _ in range(3):
                g.spmv_async(xx, yy)
            g.synchronize()
            g.set_timing("spmv")
            for _ in range(20):
                g.spmv_async(xx, yy)
            ms, cnt = g.get_timing("spmv")
            g.set_timing()
            us = ms / cnt * 1e3
            b = g.spmv_bytes(sv, sv, sched["spmv64"] if sv == 8 else sched["spmv32"])
            rec[tag] = {"us": round(us, 2), "layout_GBps": round(b / (us * 1e-6) / 1e9, 1),
                        "layout_frac": round(b / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                        "csr_GBps": round(spmv_bytes(N, sys_.nnz, sv, sv) / (us * 1e-6) / 1e9, 1)}
        g.close()
        out[name] = rec
        log(f"[bench] spmv 128^3 {name}: {rec}")
    return out


def plumbing_32():
    """The CPU-runnable case at 32^3 (SPD, shift 1e-3, x0 = 0, abs tol 1e-6, maxIter 2000):
    OpenFOAM's PCG (preconditioner none) restated on one host thread, the host ldu2csr the
    reference wrappers run per solve, and cuFoamCG through gis_foam_solve on the GPU."""
    import numpy as np
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis
    s = mesh.config_system("poisson32")
    out = {"n_cells": s.n_cells}
    try:
        from oracle import oracle as O
        _, it, res, t = O.pcg_openfoam(s, 1e-6, 2000)
        t0 = time.perf_counter()
        O.ldu2csr(s)
        out["cpu_openfoam_pcg"] = {"iterations": it, "scaled_residual": res, "s": round(t, 5)}
        out["cpu_ldu2csr_s"] = round(time.perf_counter() - t0, 5)
    except Exception as e:   # the oracle is optional at run time
        log(f"[bench] 32^3 CPU leg skipped: {e}")
    g = Gis(0)
    g.set_system(s)
    psi = np.zeros(s.n_cells)
    g.foam_solve("CG", psi, s.b, 1e-6, 2000)
    psi = np.zeros(s.n_cells)
    t0 = time.perf_counter()
    g.set_values(s.diag, s.lower, s.upper)
    perf = g.foam_solve("CG", psi, s.b, 1e-6, 2000)
    out["gpu_cuFoamCG"] = {"iterations": perf.n_iterations, "final_residual": perf.final_residual,
                           "s_with_set_values": round(time.perf_counter() - t0, 5)}
    g.close()
    log(f"[bench] 32^3 plumbing: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="mesh edge (256 = the named workload)")
    ap.add_argument("--mode", type=int, default=0, help="mixed mode: 0 reliable updates, 1 restart")
    ap.add_argument("--inner", type=float, default=1e-2,
                    help="FP64 update trigger: FP32 residual below inner * last true residual")
    ap.add_argument("--max-iter", type=int, default=100000,
                    help="iteration cap per solve (profiling runs use a small cap)")
    ap.add_argument("--no-literal", dest="literal", action="store_false",
                    help="skip the extra restarted-refinement solve with inner relTol 1e-4")
    ap.add_argument("--cpu-sample", type=int, default=30,
                    help="host CG iterations timed per one-thread variant (5x on all threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="autotune (gis_autotune) instead of keeping the schedule that "
                         "gis_set_pattern picks from the pattern (the drop-in's schedule)")
    ap.add_argument("--schedule", default=None,
                    help="fixed schedule: spmv64,spmv32,fp64_direction,fp32_direction,"
                         "coarsen64,coarsen32")
    ap.add_argument("--no-dropin", dest="dropin", action="store_false",
                    help="skip the 128^3 OpenFOAM drop-in wrapper timings")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the 64-system batched CG throughput report")
    ap.add_argument("--no-pcg", dest="pcg", action="store_false",
                    help="skip the GAMG-preconditioned CG solve of the same system")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the solver runs "
              "on a single MI355X; see the sharding notes). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis

    n = args.n
    t0 = time.perf_counter()
    s = mesh.laplacian(n, seed=0, shift=1e-4, symmetric=True)
    log(f"[bench] generated {n}^3: N={s.n_cells} nnz={s.nnz} in {time.perf_counter()-t0:.1f}s")
    N, nnz = s.n_cells, s.nnz

    g = Gis(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.set_pattern(N, s.lower_addr, s.upper_addr)
    t_pattern = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.set_values(s.diag, s.lower, s.upper)
    t_values = time.perf_counter() - t0
    t0 = time.perf_counter()
    if args.schedule:
        v = [int(t) for t in args.schedule.split(",")]
        g.set_schedule(spmv64=v[0], spmv32=v[1], fp64_direction=v[2], fp32_direction=v[3],
                       coarsen64=v[4], coarsen32=v[5])
        sched, tune_ms = g.get_schedule(), (0.0, 0.0)
    elif args.tune:
        sched, _, tune_ms = g.autotune(iterations=24)
    else:
        sched, tune_ms = g.get_schedule(), (0.0, 0.0)
    t_tune = time.perf_counter() - t0
    sched_source = "fixed" if args.schedule else ("autotune" if args.tune else "pattern")
    log(f"[bench] schedule {sched} from {sched_source} ({t_tune:.2f}s, best runs {tune_ms} ms)")
    sym_on, sym_kept = g.symmetric_storage()
    sym_state = {"in_use": sym_on, "kept_slots": sym_kept,
                 "full_slots": 64 * g.layout()["table_entries"]}
    log(f"[bench] FP64 symmetric storage {sym_state}")
    b = torch.from_numpy(s.b).cuda()
    bnorm = float(np.linalg.norm(s.b))
    max_iter = args.max_iter

    K, W = args.steps, args.warmup
    xs = [torch.zeros_like(b) for _ in range(K)]

    # ---- mixed precision (timed region; the dominant kernel is timed live) ---
    for _ in range(W):
        g.solve_mixed(b, torch.zeros_like(b), 1e-10, args.inner, max_iter, args.mode)
    torch.cuda.synchronize()
    g.set_timing("iter_spmv", period=TIMING_PERIOD)
    t0 = time.perf_counter()
    res_m, step_m = [], []
    for i in range(K):
        ts = time.perf_counter()
        res_m.append(g.solve_mixed(b, xs[i], 1e-10, args.inner, max_iter, args.mode))
        step_m.append(time.perf_counter() - ts)   # each solve ends with a host sync
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / K
    mx1_ms, mx1_n = g.get_timing("iter_spmv")
    g.set_timing()
    _, hist_m, it_m, rel_m, st_m = res_m[-1]
    log(f"[bench] mixed: {t_mixed:.3f}s/solve inner_iters={it_m} updates={len(hist_m)-1} "
        f"rel={rel_m:.3e} status={st_m}")

    # ---- plain FP64 CG (timed region) ------------------------------------------
    for _ in range(W):
        g.solve("CG", b, torch.zeros_like(b), error=1e-10 * bnorm, max_steps=max_iter,
                hist_cap=max_iter + 1)
    torch.cuda.synchronize()
    g.set_timing("iter_spmv", period=TIMING_PERIOD)
    xd = [torch.zeros_like(b) for _ in range(K)]
    t0 = time.perf_counter()
    res_d, step_d = [], []
    for i in range(K):
        ts = time.perf_counter()
        res_d.append(g.solve("CG", b, xd[i], error=1e-10 * bnorm, max_steps=max_iter,
                             hist_cap=max_iter + 1))
        step_d.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / K
    cg1_ms, cg1_n = g.get_timing("iter_spmv")
    g.set_timing()
    _, hist_d, st_d = res_d[-1]
    it_d = len(hist_d) - 1
    log(f"[bench] fp64: {t_fp64:.3f}s/solve iters={it_d} status={st_d}")

    # ---- GAMG-preconditioned CG on the same system (FP64 throughout) ---------
    pcg = None
    if args.pcg:
        t0 = time.perf_counter()
        levels = g.gamg_levels()               # builds the hierarchy (once per mesh)
        t_build = time.perf_counter() - t0
        sizes = [g.gamg_level(l)["n"] for l in range(levels)]
        def time_pcg(pc):
            for _ in range(W):
                g.solve_pcg(pc, b, torch.zeros_like(b), error=1e-10 * bnorm,
                            max_steps=max_iter, hist_cap=max_iter + 1)
            torch.cuda.synchronize()
            xp = [torch.zeros_like(b) for _ in range(K)]
            step_p = []
            for i in range(K):
                ts = time.perf_counter()
                _, hist_p, st_p = g.solve_pcg(pc, b, xp[i], error=1e-10 * bnorm,
                                              max_steps=max_iter, hist_cap=max_iter + 1)
                step_p.append(time.perf_counter() - ts)
            t_pcg = float(np.median(step_p))
            rp = torch.empty_like(b)
            g.spmv(xp[-1], rp)
            it_p = len(hist_p) - 1
            return {"solve_s": round(t_pcg, 5), "iterations": it_p,
                    "iteration_ms": round(t_pcg / max(it_p, 1) * 1e3, 3),
                    "rel_residual": float(torch.linalg.vector_norm(b - rp)) / bnorm,
                    "status": st_p}

        pcg = {"preconditioner": "GAMG", "levels": levels, "level_cells": sizes,
               "hierarchy_build_s": round(t_build, 3)}
        pcg.update(time_pcg("GAMG"))
        pcg["fp32_cycle"] = time_pcg("GAMG32")
        log(f"[bench] pcg GAMG: {pcg}")

    # ---- per-kernel breakdown: one extra solve of each with every slot timed ----
    slots = ("iter_spmv", "iter_update", "fp64_update", "iter_direction")
    g.set_timing(*slots)
    g.solve_mixed(b, torch.zeros_like(b), 1e-10, args.inner, max_iter, args.mode)
    brk_m = {s_: g.get_timing(s_) for s_ in slots}
    g.set_timing(*slots)
    g.solve("CG", b, torch.zeros_like(b), error=1e-10 * bnorm, max_steps=max_iter,
            hist_cap=max_iter + 1)
    brk_d = {s_: g.get_timing(s_) for s_ in slots}
    g.set_timing()

    def avg_us(t):
        return round(t[0] / max(t[1], 1) * 1e3, 2)

    # true residuals and mixed-vs-FP64 error, recomputed in FP64 on the device
    r = torch.empty_like(b)
    g.spmv(xs[-1], r)
    rel_true_m = float(torch.linalg.vector_norm(b - r)) / bnorm
    g.spmv(xd[-1], r)
    rel_true_d = float(torch.linalg.vector_norm(b - r)) / bnorm
    err_m = float(torch.linalg.vector_norm(xs[-1] - xd[-1]) / torch.linalg.vector_norm(xd[-1]))

    # ---- standalone SpMV bandwidth (sliced schedule, FP64 and FP32) --------
    x64 = torch.randn(N, dtype=torch.float64, device="cuda")
    y64 = torch.empty_like(x64)
    x32, y32 = x64.float(), torch.empty(N, dtype=torch.float32, device="cuda")
    spmv = {}
    for name, xx, yy, sv in (("fp64", x64, y64, 8), ("fp32", x32, y32, 4)):
        for _ in range(3):
            g.spmv_async(xx, yy)
        g.synchronize()
        g.set_timing("spmv")
        for _ in range(20):
            g.spmv_async(xx, yy)
        ms, cnt = g.get_timing("spmv")
        g.set_timing()
        us = ms / cnt * 1e3
        gbs = spmv_bytes(N, nnz, sv, sv) / (us * 1e-6) / 1e9
        lb = g.spmv_bytes(sv, sv, sched["spmv64"] if sv == 8 else sched["spmv32"])
        spmv[name] = {"us": round(us, 2), "GB/s": round(gbs, 1),
                      "frac": round(gbs / PEAK_HBM_GBS, 4),
                      "layout_bytes": lb, "layout_GBps": round(lb / (us * 1e-6) / 1e9, 1),
                      "layout_frac": round(lb / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                      "GFLOP/s": round(2 * nnz / (us * 1e-6) / 1e9, 1)}
        log(f"[bench] spmv {name}: {us:.1f} us {gbs:.0f} GB/s")

    # ---- roofline of the dominant kernel (the FP32 SpMV-carrying kernel) -----
    # The sampled launches include the no-op ones the host enqueues before it
    # sees an FP64 update raised or the solve converged (a few microseconds
    # each).  Their time is charged to the launches that work: launches timed x
    # period = all launches, the inner iterations = the launches that work.
    launches = max(mx1_n, 1) * TIMING_PERIOD
    working = sum(int(r_[2]) for r_ in res_m)
    mx1_all_s = mx1_ms / max(mx1_n, 1) / 1e3
    mx1_avg_s = mx1_all_s * launches / max(working, 1)
    split32 = sched["fp32_direction"] == 1
    kname = "k_mx1s" if split32 else "k_mx1"
    lay_b = mx1_layout_bytes(g, N, sched["spmv32"], split32)
    csr_b = mx1_bytes(N, nnz, split32)
    ach = lay_b / mx1_avg_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_poisson256.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            # counter bytes of the variant that ran (kernel form and SpMV layout)
            if tj.get(f"{kname}_spmv_schedule") == sched["spmv32"]:
                traffic = tj.get(f"{kname}_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": ("k_mx1s (FP32 SpMV of the direction + (p, q))" if split32 else
                           "k_mx1 (FP32 SpMV + direction + deferred correction)"),
                "spmv_schedule": sched["spmv32"],
                "bytes_per_launch": lay_b,
                "bytes_model": "stored matrix of the running layout (values, offset tables, "
                               "row masks) + x read once + q written once"
                               + ("" if split32 else " + r, p_old read + (every other "
                                                     "launch) xc, p_{k-2} read and xc written"),
                "achieved_csr": round(csr_b / mx1_avg_s / 1e9, 1),
                "frac_csr": round(csr_b / mx1_avg_s / 1e9 / PEAK_HBM_GBS, 4),
                "bytes_per_launch_csr": csr_b,
                "avg_launch_us": round(mx1_avg_s * 1e6, 2),
                "avg_launch_us_all": round(mx1_all_s * 1e6, 2), "launches_timed": mx1_n,
                "launches": launches, "working_launches": working,
                "avg_launch_note": "avg_launch_us: time of all launches (sampled one in "
                                   f"{TIMING_PERIOD}) over the launches that work (the inner "
                                   "iterations); avg_launch_us_all includes the no-op launches",
                "timed_one_launch_in": TIMING_PERIOD}
    if traffic:
        roofline["traffic_GBps"] = round(traffic / mx1_avg_s / 1e9, 1)
        roofline["traffic_over_bytes"] = round(traffic / lay_b, 4)
    # bytes of one FP64 CG iteration and one mixed inner iteration in the running
    # layout, for a check of the solve times against the peak.  Split form: the
    # direction pass 4.5 vector passes on average (r, p_{k-1} read, p_k written;
    # every other iteration x or xc and p_{k-2} read and x or xc written), SpMV 2,
    # update 3; the fused mixed form moves one pass less, so this bounds both
    cg_iter_b = g.spmv_bytes(8, 8, sched["spmv64"]) + 60 * N
    mx_iter_b = g.spmv_bytes(4, 4, sched["spmv32"]) + 30 * N

    literal = None
    if args.literal:
        # the configuration as stated: restarted refinement, inner relTol 1e-4, timed
        # with the same warm-up and K-step protocol as `value`
        for _ in range(W):
            g.solve_mixed(b, torch.zeros_like(b), 1e-10, 1e-4, max_iter, 1)
        torch.cuda.synchronize()
        xl = [torch.zeros_like(b) for _ in range(K)]
        step_l, res_l = [], []
        t0 = time.perf_counter()
        for i in range(K):
            ts = time.perf_counter()
            res_l.append(g.solve_mixed(b, xl[i], 1e-10, 1e-4, max_iter, 1))
            step_l.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        t_lit = (time.perf_counter() - t0) / K
        _, hl, itl, rell, stl = res_l[-1]
        literal = {"mode": "restarted refinement", "inner_rel_tol": 1e-4,
                   "s": round(t_lit, 5), "median_step_s": round(float(np.median(step_l)), 5),
                   "steps": K, "warmup": W, "iterations": itl,
                   "fp64_updates": len(hl) - 1, "rel_residual": rell, "status": stl,
                   "speedup_vs_fp64": round(t_fp64 / t_lit, 4)}
        log(f"[bench] literal restarted refinement: {literal}")

    batched = None
    if args.batch:
        batched = batch_throughput(3)
        log(f"[bench] batched: {batched}")
        multi = multi_rhs_throughput(2)
        log(f"[bench] multi-rhs: {multi}")

    drop = dropin(0 if args.no_cpu else 10) if args.dropin else None
    plumb = plumbing_32() if args.dropin else None
    unstructured = unstructured_spmv() if args.dropin else None
    pcg128 = pcg_128() if args.pcg else None
    pbicg128 = pbicg_128() if args.pcg else None

    cpu = None
    if not args.no_cpu:
        try:
            from oracle import oracle as O
            A_host = O.ldu2csr(s)
            cpu = cpu_baseline(s, A_host, xd[-1].cpu().numpy(), xs[-1].cpu().numpy(), it_d,
                               args.cpu_sample)
            del A_host
            log(f"[bench] cpu baseline: {cpu['value']:.1f}s, full solve on {cpu['cores']} "
                f"threads; GPU FP64 bit-exact: {cpu['gpu_fp64_bit_exact']}")
        except Exception as e:  # the oracle is optional at run time
            log(f"[bench] cpu baseline skipped: {e}")

    out = {
        "metric": METRIC,
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic",
        "config": {"workload": f"poisson{n}: {n}^3 7-point FV Laplacian in OpenFOAM LDU "
                               f"layout, off-diagonals -(1+0.5u), shift 1e-4, seed 0, "
                               f"b~U[-1,1], x0=0, CG to ||b-Ax||/||b|| <= 1e-10",
                   "n_cells": N, "nnz": nnz,
                   "mixed_mode": "reliable updates" if args.mode == 0 else "restarted refinement",
                   "inner_rel_tol": args.inner},
        "fp64_cg_s": round(t_fp64, 5),
        "median_step_s": {"mixed": round(float(np.median(step_m)), 5),
                          "fp64": round(float(np.median(step_d)), 5)},
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 4),
        "iterations_fp64": it_d,
        "iterations_mixed": it_m,
        "fp64_updates_mixed": len(hist_m) - 1,
        "rel_residual_mixed": rel_true_m,
        "rel_residual_fp64": rel_true_d,
        "rel_error_mixed_vs_fp64": err_m,
        "rel_error_mixed_vs_oracle_fp64": (cpu or {}).get("rel_error_mixed_vs_oracle_fp64"),
        "layout_bytes_per_iteration": {"fp64_cg": cg_iter_b, "mixed_inner": mx_iter_b,
                                       "fp64_cg_TBps": round(cg_iter_b * it_d / t_fp64 / 1e12, 3),
                                       "mixed_TBps": round(mx_iter_b * it_m / t_mixed / 1e12, 3)},
        "fp64_iteration_us": round(t_fp64 / max(it_d, 1) * 1e6, 2),
        "mixed_iteration_us": round(t_mixed / max(it_m, 1) * 1e6, 2),
        "kernel_avg_us": {"mixed_spmv_kernel": avg_us((mx1_ms, mx1_n)),
                          "fp64_spmv_kernel": avg_us((cg1_ms, cg1_n)),
                          "instrumented_solve": {
                              "mixed": {"spmv_kernel": avg_us(brk_m["iter_spmv"]),
                                        "update_k_mx2": avg_us(brk_m["iter_update"]),
                                        "fp64_update_k_mxu": avg_us(brk_m["fp64_update"]),
                                        "direction_k_mx_dir": avg_us(brk_m["iter_direction"])},
                              "fp64": {"spmv_kernel": avg_us(brk_d["iter_spmv"]),
                                       "update_k_cg2": avg_us(brk_d["iter_update"]),
                                       "direction_k_cg_dir": avg_us(brk_d["iter_direction"])}}},
        "schedule": sched,
        "schedule_source": sched_source,
        "symmetric_storage_fp64": sym_state,
        "spmv": spmv,
        "setup_s": {"pattern": round(t_pattern, 4), "values": round(t_values, 4),
                    "autotune": round(t_tune, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if literal is not None:
        out["mixed_restarted_inner_1e-4"] = literal
    if drop is not None:
        out["dropin_128"] = drop
        out["plumbing_32"] = plumb
        out["spmv_128_numberings"] = unstructured
    if batched is not None:
        out["batched_cg_64_systems"] = batched
        out["spmm_bicg_3_rhs"] = multi
    if pcg is not None:
        for rec in (pcg, pcg["fp32_cycle"]):
            rec["speedup_vs_fp64_cg"] = round(t_fp64 / rec["solve_s"], 3)
            rec["speedup_vs_mixed"] = round(t_mixed / rec["solve_s"], 3)
        out["pcg_gamg_to_1e-10"] = pcg
    if pcg128 is not None:
        out["pcg_128_preconditioners"] = pcg128
    if pbicg128 is not None:
        out["pbicg_128_preconditioners"] = pbicg128
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
