"""A compiled C++ caller of the C ABI (tests/cabi/foam_sequence.cpp): it
includes only include/gis.h and the OpenFOAM wrappers' bridge header, is built
with g++ and no ROCm include path, and links libgis_amd.so.

CPU: the header compiles without ROCm, and the struct layouts it sees equal
the ctypes mirrors in cufoam_amd/gis.py.  GPU: the cuFoam<Name>::solve sequence
(handle per addressing, gis_set_pattern once per mesh, gis_set_values and
gis_foam_solve per solve) against the oracle's restatement of the wrappers
(cuFoamCG.C:38-89 and siblings), bit for bit."""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from cufoam_amd import gis as G
from cufoam_amd import mesh

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cabi"))
import build_caller  # noqa: E402


def _offsets(struct, names):
    return [C.sizeof(struct)] + [getattr(struct, n).offset for n in names]


def test_header_layouts_match_the_ctypes_mirrors():
    exe = build_caller.build()
    out = subprocess.run([exe, "layout"], check=True, capture_output=True, text=True).stdout
    got = json.loads(out)
    assert got["gis_params"] == _offsets(G._Params, [f[0] for f in G._Params._fields_])
    assert got["gis_performance"] == _offsets(G._Perf, [f[0] for f in G._Perf._fields_])
    assert got["gis_schedule"] == _offsets(G._Schedule, [f[0] for f in G._Schedule._fields_])


def _code(method):
    """foam_sequence's method code: a GIS_ id, 100 + p for PCG, 200 + p for PBiCG."""
    if ":" in method:
        family, pc = method.split(":")
        return (100 if family == "PCG" else 200) + G.PRECONDITIONERS[pc]
    return G.METHODS[method]


def _write_systems(path, cases):
    with open(path, "wb") as f:
        np.array([len(cases)], np.int32).tofile(f)
        for method, s, psi, tol, max_iter in cases:
            np.array([_code(method), s.n_cells, s.n_faces, max_iter], np.int32).tofile(f)
            np.array([tol], np.float64).tofile(f)
            for a, dt in ((s.lower_addr, np.int32), (s.upper_addr, np.int32),
                          (s.diag, np.float64), (s.lower, np.float64), (s.upper, np.float64),
                          (s.b, np.float64), (psi, np.float64)):
                np.ascontiguousarray(a, dt).tofile(f)


@pytest.mark.gpu
def test_compiled_caller_runs_the_wrapper_sequence_bit_exact(tmp_path):
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    cases = []
    for method, s, tol, max_iter in (
            ("CG", mesh.laplacian(32, seed=0, shift=1e-3), 1e-6, 2000),
            ("BiCG", mesh.laplacian(16, seed=1, shift=0.1, symmetric=False), 1e-8, 2000),
            ("MRJacobi", mesh.random_ldu(3000, 9000, seed=2, symmetric=False), 1e-8, 3000),
            ("Richardson", mesh.laplacian(9, seed=3, shift=0.5), 1e-8, 500),
            ("JMR", mesh.laplacian(11, seed=4, shift=0.3, symmetric=False), 1e-8, 2000),
            ("PBiCG:diagonal", mesh.random_ldu(2000, 7000, seed=5, symmetric=False), 1e-8,
             2000),
            ("PCG:DIC", mesh.laplacian(12, seed=6, shift=1e-2), 1e-8, 2000),
            ("PBiCG:DILU", mesh.laplacian(10, seed=7, shift=1e-2, symmetric=False), 1e-8,
             2000)):
        cases.append((method, s, rng.uniform(-1, 1, s.n_cells), tol, max_iter))
    exe = build_caller.build()
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_systems(src, cases)
    print(f"[case] compiled C++ caller, {len(cases)} systems", flush=True)
    r = subprocess.run([exe, "solve", str(src), str(dst)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    with open(dst, "rb") as f:
        builds = int(np.fromfile(f, np.int64, 1)[0])
        assert builds == len(cases)   # one pattern per new mesh, none for the repeat solve
        for method, s, psi, tol, max_iter in cases:
            st, iters, same = np.fromfile(f, np.int32, 3)
            init, fin, nf = np.fromfile(f, np.float64, 3)
            got = np.fromfile(f, np.float64, s.n_cells)
            family, _, pc = method.partition(":")
            ref_x, ref = O.foam_solve(family, s, psi, s.b, tol, max_iter,
                                      preconditioner=pc or "none")
            assert same == 1, method
            assert (iters, init, fin) == (ref.n_iterations, ref.initial_residual,
                                          ref.final_residual), method
            assert nf == O.norm_factor(O.ldu2csr(s), s.b, psi)
            np.testing.assert_array_equal(got, ref_x)
