"""The allocation-free NormFactor entry point (gis_norm_factor_device_noalloc)
against gis_norm_factor_device and the oracle, bit for bit, through the C ABI.

Its workspace is one fixed device array per GPU shared by every call, so the
tests run calls of different sizes back to back (a larger tree after a smaller
one must refill the slots first), recover from a reduction that gave up, and
check the size limit on the host.
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 25


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda:0")


def _check_case(g, n, seed, shift):
    print(f"[case] noalloc NormFactor n={n} seed={seed}", flush=True)
    s = mesh.laplacian(n, seed=seed, shift=shift, symmetric=(seed % 2 == 0))
    g.set_system(s)
    A = O.ldu2csr(s)
    x = np.random.default_rng(seed + 100).uniform(-1, 1, s.n_cells)
    for xv in (x, np.zeros(s.n_cells)):
        b_d, x_d = _dev(s.b), _dev(xv)
        got = g.norm_factor_noalloc(b_d, x_d)
        assert got == g.norm_factor(b_d, x_d)
        assert got == O.norm_factor(A, s.b, xv)


def test_noalloc_matches_allocating_entry_two_calls(gpu):
    # 100^3: 977 tiles (16 groups, level-1 collector); then 24^3: 14 tiles
    _check_case(gpu, 100, 0, 1e-3)
    _check_case(gpu, 24, 1, 0.1)


def test_noalloc_larger_tree_after_smaller(gpu):
    # 3 tiles, then 64^3 (256 tiles: 4 groups) and 1 tile: the prefix of empty
    # slots grows and shrinks between calls on the shared array
    _check_case(gpu, 14, 2, 1e-3)
    _check_case(gpu, 64, 3, 1e-2)
    _check_case(gpu, 1, 4, 1.0)


def test_noalloc_after_failed_reduction(gpu):
    from cufoam_amd.gis import GisError
    s = mesh.laplacian(48, seed=5, shift=1e-2)   # 108 tiles, 2 groups
    gpu.set_system(s)
    b_d, x_d = _dev(s.b), _dev(np.ones(s.n_cells))
    gpu.set_slot_fault(4)
    try:
        print("[case] noalloc NormFactor with the slot-fault hook", flush=True)
        with pytest.raises(GisError) as e:
            gpu.norm_factor_noalloc(b_d, x_d)
        assert e.value.status == 13
        with pytest.raises(GisError) as e:
            gpu.norm_factor(b_d, x_d)
        assert e.value.status == 13
    finally:
        gpu.set_slot_fault(0)
    A = O.ldu2csr(s)
    ref = O.norm_factor(A, s.b, np.ones(s.n_cells))
    assert gpu.norm_factor_noalloc(b_d, x_d) == ref
    assert gpu.norm_factor(b_d, x_d) == ref
    _check_case(gpu, 32, 6, 1e-3)


def test_noalloc_rejects_more_rows_than_its_array():
    from cufoam_amd.gis import Gis, GisError
    print("[case] noalloc NormFactor size limit", flush=True)
    n = MAX_ROWS + 1
    with Gis(0) as g:
        g.set_pattern(n, np.zeros(0, np.int32), np.zeros(0, np.int32))
        g.set_values(np.full(n, 2.0), np.zeros(0), np.zeros(0))
        b_d = _dev(np.ones(n))
        with pytest.raises(GisError) as e:
            g.norm_factor_noalloc(b_d, b_d)
        assert e.value.status == 3
        # the allocating entry point takes the same system
        assert g.norm_factor(b_d, b_d) > 0.0
