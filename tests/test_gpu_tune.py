"""GPU: the schedule autotuner.  Every candidate it may choose is bit-identical
(SpMV parity gate inside the tuner), the choice is applied and cached per
(device, N, nnz), and solves after tuning still match the oracle bit for bit."""
import numpy as np
import pytest

from cufoam_amd import mesh
from cufoam_amd.gis import Gis, GisError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_autotune_applies_caches_and_keeps_parity(gpu):
    Gis.autotune_clear()
    s = mesh.laplacian(24, seed=3, shift=1e-3)
    print("[case] autotune 24^3", flush=True)
    gpu.set_system(s)
    sched, cached, ms = gpu.autotune(iterations=12)
    assert not cached and ms[0] > 0 and ms[1] > 0
    assert sched["spmv64"] in range(7) and sched["fp64_direction"] in (0, 1)
    assert sched["spmv32"] in range(7) and sched["fp32_direction"] in (0, 1)
    assert gpu.get_schedule() == sched
    # a second handle on the same mesh takes the cached choice without timing
    with Gis(0) as other:
        other.set_system(s)
        sched2, cached2, ms2 = other.autotune(iterations=12)
        assert cached2 and sched2 == sched and ms2 == (0.0, 0.0)
    # solves under the tuned schedule reproduce the oracle
    A = O.ldu2csr(s)
    err = 1e-8 * np.linalg.norm(s.b)
    ref = O.solve("CG", A, s.b, s.x0, err, 2000)
    x, hist, st = gpu.solve("CG", s.b, s.x0, error=err, max_steps=2000)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
    rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 5000, 0)
    xm, hm, it, rel, stm = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 5000, 0)
    assert (stm, it) == (rst, rit)
    np.testing.assert_array_equal(xm, rx)
    Gis.autotune_clear()
    _, cached3, _ = gpu.autotune(fp32=False, iterations=4)
    assert not cached3


def test_autotune_rejects_bad_targets(gpu):
    s = mesh.laplacian(4, seed=0)
    gpu.set_system(s)
    with pytest.raises(GisError):
        gpu.autotune(fp64=False, fp32=False)


def test_pattern_picks_its_schedule_without_timing(gpu):
    """gis_set_pattern starts a structured mesh (every 64-row slice tabled) on
    SellD (FP64 with the separate direction pass from 2^19 rows up, the
    direction formed at the gathers below and in FP32), an unstructured one on
    sliced rows; gis_get_layout reports the counts the SellD byte model uses."""
    from cufoam_amd import mesh
    from cufoam_amd.gis import Gis
    Gis.autotune_clear()
    s = mesh.laplacian(32, seed=0, shift=1e-3)
    gpu.set_system(s)
    sch = gpu.get_schedule()
    assert (sch["spmv64"], sch["spmv32"], sch["fp64_direction"], sch["fp32_direction"]) == (6, 6, 0, 0)
    lay = gpu.layout()
    assert lay["n_rows"] == 32768 and lay["nnz"] == s.nnz and lay["slices"] == 512
    assert lay["table_slices"] == 512 and lay["untabled_slots"] == 0
    assert lay["byte_mask_slices"] == 512
    # a slice lies on one x-line: offsets -n^2, -n, -1, 0, +1, +n, +n^2 minus the
    # missing neighbours of boundary lines
    assert 5 * 512 <= lay["table_entries"] <= 7 * 512
    csr = s.nnz * 12 + 4 * (s.n_cells + 1) + 16 * s.n_cells
    assert gpu.spmv_bytes(8, 8, 6) < csr
    big = mesh.laplacian(81, seed=0, shift=1e-3)   # 531,441 rows
    gpu.set_system(big)
    assert gpu.get_schedule()["fp64_direction"] == 1
    r = mesh.random_ldu(5000, 15000, seed=1)
    gpu.set_system(r)
    assert gpu.get_schedule()["spmv64"] == 0
    lay = gpu.layout()
    assert lay["table_slices"] + (lay["untabled_slots"] > 0) >= 1
    assert lay["sell_slots"] >= r.nnz

