"""CPU checks of the C ABI: the library builds for gfx950, loads, and exports
every symbol include/gis.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

from cufoam_amd import build as B
from cufoam_amd import gis as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "gis.h")).read()
    return sorted(set(re.findall(r"GIS_API\s+[\w\s\*]+?\b(gis_\w+)\s*\(", hdr)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ["gis_create", "gis_destroy", "gis_set_pattern", "gis_set_values",
                 "gis_get_csr", "gis_solve", "gis_foam_solve", "gis_norm_factor"]:
        assert must in syms


def test_library_loads_and_exports_all_symbols():
    if not os.path.exists(B.LIB):
        B.build()
    lib = ctypes.CDLL(B.LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"symbols declared in gis.h but not exported: {missing}"


def test_python_binding_covers_header():
    assert set(declared_symbols()) <= set(G.SIGNATURES)


def test_status_strings_without_gpu():
    lib = G.load()
    assert lib.gis_status_string(0) == b"ok"
    assert lib.gis_status_string(10) == b"BiCG breakdown"


def test_default_params_match_reference_defaults():
    # gis/gis.h:13-87: error 1e-4, max_steps 0, omega 0, power steps 30, int_bound 1e-10
    lib = G.load()
    p = G._Params()
    lib.gis_default_params(ctypes.byref(p))
    p.detail_type = 7
    lib.gis_default_params(ctypes.byref(p))
    assert (p.error, p.max_steps, p.omega, p.power_steps, p.int_bound) == (1e-4, 0, 0.0, 30, 1e-10)
    assert p.detail_type == 0   # DetailType::Error, the reference's default


def test_create_without_gpu_reports_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lib = G.load()
    h = ctypes.c_void_p()
    assert lib.gis_create(ctypes.byref(h), 0) != 0


def test_foam_convergence_control_is_validated_before_any_gpu_call():
    from cufoam_amd import foam, mesh
    s = mesh.laplacian(4, seed=0, shift=0.1)
    m = foam.lduMatrix.from_system(s)
    with pytest.raises(foam.FatalError):
        foam.New("cuFoamCG", "p", m, {"convergence": "strict"})


def test_foam_pbicg_registration_and_preconditioner_checks_before_any_gpu_call():
    from cufoam_amd import foam, mesh
    asym = foam.lduMatrix.from_system(mesh.laplacian(4, seed=0, shift=0.1, symmetric=False))
    assert asym.asymmetric()
    with pytest.raises(foam.FatalError):   # PCG is registered for symmetric matrices only
        foam.New("cuFoamPCG", "U", asym, {"preconditioner": "diagonal"})
    for pc in ("DIC", "GAMG", "GAMG32"):   # PBiCG takes none, diagonal or DILU
        with pytest.raises(foam.FatalError):
            foam.New("cuFoamPBiCG", "U", asym, {"preconditioner": pc})
    sym = foam.lduMatrix.from_system(mesh.laplacian(4, seed=0, shift=0.1))
    with pytest.raises(foam.FatalError):   # DILU is for the asymmetric solver
        foam.New("cuFoamPCG", "p", sym, {"preconditioner": "DILU"})
