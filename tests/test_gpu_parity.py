"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer work (pattern, permutation) and every FP64 result must be bit-exact:
the kernels and the oracle share one arithmetic contract (explicit fma,
sequential CSR row sums, one pairwise reduction order fixed by the length).
At full workload sizes, size-independent properties stand in for the oracle.
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SMALL = [(1, 0), (2, 1), (4, 0), (8, 3), (16, 2), (32, 0)]


def _torch():
    import torch
    return torch


def _case(n, seed, symmetric=True, shift=1e-3):
    print(f"[case] n={n} seed={seed} symmetric={symmetric}", flush=True)
    return mesh.laplacian(n, seed=seed, shift=shift, symmetric=symmetric)


@pytest.mark.parametrize("n,seed", SMALL)
@pytest.mark.parametrize("symmetric", [True, False])
def test_ldu2csr_bit_exact(gpu, n, seed, symmetric):
    s = _case(n, seed, symmetric, shift=0.1)
    gpu.set_system(s)
    rp, col, val = gpu.get_csr()
    A = O.ldu2csr(s)
    assert np.array_equal(rp, A.rowptr)
    assert np.array_equal(col, A.col)
    assert np.array_equal(val.view(np.uint64), A.val.view(np.uint64))
    rpt, colt, valt = gpu.get_csr(transpose=True)
    At = O.transpose(A)
    assert np.array_equal(rpt, At.rowptr) and np.array_equal(colt, At.col)
    assert np.array_equal(valt.view(np.uint64), At.val.view(np.uint64))


def test_ldu2csr_bit_exact_128(gpu):
    s = _case(128, 0, symmetric=False, shift=0.1)
    gpu.set_system(s)
    rp, col, val = gpu.get_csr()
    A = O.ldu2csr(s)
    assert np.array_equal(rp, A.rowptr)
    assert np.array_equal(col, A.col)
    assert np.array_equal(val.view(np.uint64), A.val.view(np.uint64))


def test_ldu2csr_rejects_unsorted_owner(gpu):
    from cufoam_amd.gis import GisError
    la = np.array([1, 0], np.int32)
    ua = np.array([2, 2], np.int32)
    with pytest.raises(GisError):
        gpu.set_pattern(3, la, ua)


def _hub(n):
    """Cell 0 coupled to every other cell: row 0 holds n entries, the others 2."""
    return np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32)


def test_hub_pattern_bit_exact(gpu):
    torch = _torch()
    n = 5000
    la, ua = _hub(n)
    rng = np.random.default_rng(11)
    upper = -rng.uniform(0.5, 1.0, n - 1)
    lower = upper * rng.uniform(1.0, 1.5, n - 1)
    diag = np.full(n, 1.0)
    diag[0] += np.abs(lower).sum()
    diag[1:] += np.abs(upper) + 0.1
    s = mesh.LduSystem(n, la, ua, diag, lower, upper, np.ones(n), np.zeros(n), meta={})
    print("[case] hub pattern n=5000", flush=True)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    rp, col, val = gpu.get_csr()
    assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
    assert np.array_equal(val, A.val)
    x = rng.uniform(-1, 1, n)
    for sched in range(7):
        gpu.set_schedule(spmv64=sched)
        y = gpu.spmv(torch.as_tensor(x, device="cuda:0")).cpu().numpy()
        assert np.array_equal(y, O.spmv(A, x)), sched
    gpu.set_schedule(spmv64=0)


def test_sliced_layout_past_2_31_slots_is_rejected():
    """A hub row of 2^25 entries pads slice 0 to 2^31 slots (nnz is only 1e8):
    set_pattern reports GIS_ERR_SHAPE instead of overflowing the slot offsets."""
    from cufoam_amd.gis import Gis, GisError
    n = 1 << 25
    la, ua = _hub(n)
    print("[case] hub pattern n=2^25 (slot overflow)", flush=True)
    with Gis(0) as g:
        with pytest.raises(GisError) as e:
            g.set_pattern(n, la, ua)
        assert e.value.status == 3
        g.set_pattern(3, np.array([0, 1], np.int32), np.array([1, 2], np.int32))
        assert g.n == 3


@pytest.mark.parametrize("n,seed", SMALL + [(64, 5)])
def test_spmv_dot_bit_exact(gpu, n, seed):
    torch = _torch()
    s = _case(n, seed, symmetric=False, shift=0.1)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    x = np.random.default_rng(seed + 7).standard_normal(s.n_cells)
    dx = torch.from_numpy(x).cuda()
    y = gpu.spmv(dx).cpu().numpy()
    assert np.array_equal(y, O.spmv(A, x))
    yt = gpu.spmv(dx, transpose=True).cpu().numpy()
    assert np.array_equal(yt, O.spmv(O.transpose(A), x))
    d = gpu.dot(dx, torch.from_numpy(y).cuda())
    assert d == O.dot(x, y)


def _mixed_locality_system(seed):
    """A 16^3 hex block (compressible slot columns) followed by a 4096-cell block
    with random faces (offsets far beyond a byte): both SpMV column paths."""
    a = mesh.laplacian(16, seed=seed, shift=0.1, symmetric=False)
    r = mesh.random_ldu(4096, 12000, seed=seed, symmetric=False)
    n0 = a.n_cells
    la = np.concatenate([a.lower_addr, r.lower_addr + n0]).astype(np.int32)
    ua = np.concatenate([a.upper_addr, r.upper_addr + n0]).astype(np.int32)
    return mesh.LduSystem(n0 + r.n_cells, la, ua, np.concatenate([a.diag, r.diag]),
                          np.concatenate([a.lower, r.lower]), np.concatenate([a.upper, r.upper]),
                          np.concatenate([a.b, r.b]), np.zeros(n0 + r.n_cells))


@pytest.mark.parametrize("schedule", [0, 4, 5, 6])
def test_spmv_schedules_mixed_locality(gpu, schedule):
    torch = _torch()
    s = _mixed_locality_system(3)
    print(f"[case] mixed locality N={s.n_cells} schedule={schedule}", flush=True)
    gpu.set_system(s)
    gpu.set_schedule(schedule, schedule, -1)
    A = O.ldu2csr(s)
    x = np.random.default_rng(9).standard_normal(s.n_cells)
    dx = torch.from_numpy(x).cuda()
    assert np.array_equal(gpu.spmv(dx).cpu().numpy(), O.spmv(A, x))
    assert np.array_equal(gpu.spmv(dx, transpose=True).cpu().numpy(),
                          O.spmv(O.transpose(A), x))
    x32 = dx.float()
    y32 = torch.empty_like(x32)
    gpu.spmv_async(x32, y32)
    gpu.synchronize()
    assert np.array_equal(y32.cpu().numpy(), O.spmv_f32(A, x.astype(np.float32)))
    ref = O.solve("BiCG", A, s.b, s.x0, 0.0, 30)
    xs, hist, st = gpu.solve("BiCG", s.b, s.x0, error=0.0, max_steps=30)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(xs, ref.x)


@pytest.mark.parametrize("n,seed", SMALL)
def test_norm_factor_bit_exact(gpu, n, seed):
    s = _case(n, seed)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    x = np.random.default_rng(seed).uniform(-2, 2, s.n_cells)
    assert gpu.norm_factor(s.b, x) == O.norm_factor(A, s.b, x)
    assert gpu.norm_factor(s.b, s.x0) == O.norm_factor(A, s.b, s.x0)


@pytest.mark.parametrize("n,seed", SMALL)
@pytest.mark.parametrize("max_steps", [1, 7, 0])
def test_cg_bit_exact(gpu, n, seed, max_steps):
    s = _case(n, seed)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    x0 = np.random.default_rng(seed + 1).uniform(-1, 1, s.n_cells) if seed % 2 else s.x0
    err = 1e-9 * np.linalg.norm(s.b)
    ref = O.solve("CG", A, s.b, x0, err, max_steps)
    x, hist, st = gpu.solve("CG", s.b, x0, error=err, max_steps=max_steps,
                            hist_cap=ref.hist_len + 8)
    assert len(hist) == ref.hist_len
    assert np.array_equal(hist, ref.hist)
    assert np.array_equal(x, ref.x)


def test_cg_device_tensors_128(gpu):
    torch = _torch()
    s = _case(128, 0)
    gpu.set_system(s)
    b = torch.from_numpy(s.b).cuda()
    x = torch.zeros_like(b)
    err = 1e-8 * float(np.linalg.norm(s.b))
    x, hist, st = gpu.solve("CG", b, x, error=err, max_steps=3000)
    assert hist[-1] <= err < hist[-2]
    A = O.ldu2csr(s)
    r = s.b - O.spmv(A, x.cpu().numpy())
    # recursive and true residual agree to rounding level
    assert abs(np.linalg.norm(r) - hist[-1]) <= 1e-6 * np.linalg.norm(s.b)


@pytest.mark.parametrize("n,seed", [(8, 0), (16, 1), (32, 2)])
def test_foam_cg_matches_oracle_wrapper(gpu, n, seed):
    from cufoam_amd import foam
    s = _case(n, seed)
    psi = np.zeros(s.n_cells)
    m = foam.lduMatrix.from_system(s)
    solver = foam.New("cuFoamCG", "p", m, {"tolerance": 1e-7, "maxIter": 500}, gis=gpu)
    perf = solver.solve(psi, s.b)
    ref_x, ref = O.foam_solve("CG", s, np.zeros(s.n_cells), s.b, 1e-7, 500)
    assert perf.n_iterations == ref.n_iterations
    assert perf.initial_residual == ref.initial_residual
    assert perf.final_residual == ref.final_residual
    assert np.array_equal(psi, ref_x)


def test_kernel_timing_samples_one_launch_in_period(gpu):
    s = mesh.laplacian(24, seed=31, shift=1e-2)
    gpu.set_system(s)
    for period, expect in ((1, 40), (4, 10)):
        gpu.set_timing("iter_spmv", period=period)
        gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=40)
        ms, n = gpu.get_timing("iter_spmv")
        gpu.set_timing()
        assert n == expect and ms > 0.0
