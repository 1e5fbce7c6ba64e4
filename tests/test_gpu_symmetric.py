"""GPU parity of the symmetric storage of the offset-table SpMV schedule
(gis_set_symmetric_storage, gis_get_symmetric_storage; no counterpart in the
reference).  Where every 64-row slice has a column-offset table and the values
are bitwise symmetric, the schedule keeps only the slots on and above the
diagonal and reads each slot below it from its mirror.  Checked here: it is in
use exactly when it may be (symmetric values, every slice tabled, switched on),
it follows gis_set_values (one changed bit, -0.0 against +0.0, and back), and
SpMV (FP64, FP32), CG, mixed-precision CG and PCG results equal the oracle's
and those of the full storage bit for bit, on 7-point and 27-point stencils."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _lap(n, seed, shift=1e-3):
    print(f"[case] symmetric laplacian n={n} seed={seed}", flush=True)
    return mesh.laplacian(n, seed=seed, shift=shift)


def _spmv_all(gpu, s, A, seed):
    import torch
    x = np.random.default_rng(seed).standard_normal(s.n_cells)
    dx = torch.from_numpy(x).cuda()
    y64 = gpu.spmv(dx).cpu().numpy()
    x32 = dx.float()
    y32 = torch.empty_like(x32)
    gpu.spmv_async(x32, y32)
    gpu.synchronize()
    assert np.array_equal(y64, O.spmv(A, x))
    assert np.array_equal(y32.cpu().numpy(), O.spmv_f32(A, x.astype(np.float32)))
    return y64, y32.cpu().numpy()


# 35^3 = 669 x 64 + 59 rows: a last slice with a table and 5 rows past N
@pytest.mark.parametrize("n,seed", [(8, 1), (16, 2), (35, 4), (40, 3)])
def test_spmv_symmetric_storage_bit_exact(gpu, n, seed):
    s = _lap(n, seed)
    gpu.set_system(s)
    gpu.set_schedule(6, 6, -1)
    try:
        A = O.ldu2csr(s)
        lay = gpu.layout()
        assert lay["table_slices"] == lay["slices"]
        active, kept = gpu.symmetric_storage()
        full = 64 * gpu.layout()["table_entries"]
        # a 7-point stencil keeps the diagonal and the upper offsets: 4 of 7 inside
        assert active and 3 * full < 7 * kept < 5 * full
        y_sym = _spmv_all(gpu, s, A, seed + 10)
        gpu.set_symmetric_storage(False)
        assert gpu.symmetric_storage()[0] is False
        y_full = _spmv_all(gpu, s, A, seed + 10)
        for a, b in zip(y_sym, y_full):
            assert np.array_equal(a, b)
    finally:
        gpu.set_symmetric_storage(True)
        gpu.set_schedule(0, 0, -1)


def test_solvers_symmetric_storage_bit_exact(gpu):
    s = _lap(32, 4, shift=1e-4)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    gpu.set_schedule(6, 6, -1)
    try:
        assert gpu.symmetric_storage()[0]
        ref = O.solve("CG", A, s.b, s.x0, 0.0, 30)
        rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 5000, 0)
        results = []
        for on in (True, False):
            gpu.set_symmetric_storage(on)
            assert gpu.symmetric_storage()[0] is on
            for direction in (0, 1):
                gpu.set_schedule(6, 6, -1, direction, direction, -1, -1)
                x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=30, hist_cap=40)
                np.testing.assert_array_equal(hist, ref.hist)
                np.testing.assert_array_equal(x, ref.x)
                xm, hm, it, rel, stm = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 5000, 0)
                assert (stm, it) == (rst, rit)
                np.testing.assert_array_equal(hm, rh)
                np.testing.assert_array_equal(xm, rx)
            xp, hp, stp = gpu.solve_pcg("DIC", s.b, s.x0, error=0.0, max_steps=12, hist_cap=20)
            xg, hg, stg = gpu.solve_pcg("GAMG", s.b, s.x0, error=0.0, max_steps=8, hist_cap=20)
            results.append((xp, hp, xg, hg))
        for a, b in zip(results[0], results[1]):
            np.testing.assert_array_equal(a, b)
        ref_dic = O.pcg("DIC", s, s.b, s.x0, 0.0, 12)
        ref_gamg = O.pcg("GAMG", s, s.b, s.x0, 0.0, 8)
        np.testing.assert_array_equal(results[0][0], ref_dic.x)
        np.testing.assert_array_equal(results[0][1], ref_dic.hist)
        np.testing.assert_array_equal(results[0][2], ref_gamg.x)
        np.testing.assert_array_equal(results[0][3], ref_gamg.hist)
    finally:
        gpu.set_symmetric_storage(True)
        gpu.set_schedule(0, 0, -1, 0, 0, 1, 1)


def test_symmetric_storage_follows_values(gpu):
    s = _lap(16, 5)
    gpu.set_system(s)
    gpu.set_schedule(6, 6, -1)
    try:
        assert gpu.symmetric_storage()[0]
        # one lower coefficient one ulp away: not symmetric, full storage
        lower = s.lower.copy()
        lower[len(lower) // 2] = np.nextafter(lower[len(lower) // 2], 0.0)
        t = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, lower, s.upper,
                           s.b, s.x0)
        gpu.set_values(t.diag, t.lower, t.upper)
        assert gpu.symmetric_storage()[0] is False
        A = O.ldu2csr(t)
        _spmv_all(gpu, t, A, 51)
        ref = O.solve("CG", A, t.b, t.x0, 0.0, 20)
        x, hist, st = gpu.solve("CG", t.b, t.x0, error=0.0, max_steps=20, hist_cap=24)
        np.testing.assert_array_equal(hist, ref.hist)
        np.testing.assert_array_equal(x, ref.x)
        # -0.0 below and +0.0 above the diagonal: equal numbers, different bits
        up = s.upper.copy()
        lo = s.lower.copy()
        up[3], lo[3] = 0.0, -0.0
        z = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, lo, up, s.b, s.x0)
        gpu.set_values(z.diag, z.lower, z.upper)
        assert gpu.symmetric_storage()[0] is False
        _spmv_all(gpu, z, O.ldu2csr(z), 52)
        # symmetric again
        gpu.set_values(s.diag, s.lower, s.upper)
        assert gpu.symmetric_storage()[0]
        _spmv_all(gpu, s, O.ldu2csr(s), 53)
    finally:
        gpu.set_schedule(0, 0, -1)


def test_symmetric_storage_needs_every_slice_tabled(gpu):
    # 33^3 = 35937 rows: the last slice holds 33 rows and pads beyond a quarter
    s = _lap(33, 6)
    gpu.set_system(s)
    lay = gpu.layout()
    assert lay["table_slices"] < lay["slices"]
    assert gpu.symmetric_storage() == (False, 0)
    # nonsymmetric values on a fully tabled pattern: layout kept, not in use
    t = mesh.laplacian(16, seed=7, shift=0.1, symmetric=False)
    gpu.set_system(t)
    active, kept = gpu.symmetric_storage()
    assert not active and kept > 0


def _masks_system(seed):
    """The 7-point 12^3 + 27-point 16^3 system of test_gpu_selld_masks.py."""
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(__file__), "test_gpu_selld_masks.py")
    spec = importlib.util.spec_from_file_location("_selld_masks_cases", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod._system(seed)


def test_symmetric_storage_27_point(gpu):
    s = _masks_system(31)
    print("[case] 7-point 12^3 + 27-point 16^3 seed=31", flush=True)
    gpu.set_system(s)
    lay = gpu.layout()
    gpu.set_schedule(6, 6, -1)
    try:
        A = O.ldu2csr(s)
        if lay["table_slices"] < lay["slices"]:
            pytest.skip("not every slice has a table")
        assert gpu.symmetric_storage()[0]
        _spmv_all(gpu, s, A, 61)
        ref = O.solve("CG", A, s.b, s.x0, 0.0, 25)
        x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=25, hist_cap=32)
        np.testing.assert_array_equal(hist, ref.hist)
        np.testing.assert_array_equal(x, ref.x)
    finally:
        gpu.set_schedule(0, 0, -1)


def test_symmetric_storage_argument_checks(gpu):
    from cufoam_amd.gis import GisError
    with pytest.raises(GisError) as e:
        gpu._check(gpu._L.gis_set_symmetric_storage(gpu._h, 2), "set_symmetric_storage")
    assert e.value.status == 1
