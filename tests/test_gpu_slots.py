"""GPU parity of the slot reductions where every level of the collecting tree is
used: at 170^3 (4,913,000 rows, 4,798 tiles) the tree has 4,798 tile partials,
75 group sums, 2 level-2 sums and the total, all of them ragged, so group
collectors, the top collector's level-1 pass over several waves and its LDS
level-2 fold all run.  Dot, NormFactor (two totals), CG and mixed-precision CG
(one total per kernel, several kernels sharing the slots) are compared bit for
bit with the oracle, twice in a row so that the restored slots are exercised."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_EDGE = 170


@pytest.fixture(scope="module")
def big():
    print(f"[case] laplacian {N_EDGE}^3 seed=21", flush=True)
    s = mesh.laplacian(N_EDGE, seed=21, shift=1e-3)
    return s, O.ldu2csr(s)


def test_tree_shape_uses_every_collector_level():
    tiles = -(-N_EDGE ** 3 // 1024)
    counts = [tiles]
    while counts[-1] > 1:
        counts.append(-(-counts[-1] // 64))
    assert counts == [4798, 75, 2, 1]


def test_dot_and_norm_factor_bit_exact(gpu, big):
    s, A = big
    gpu.set_system(s)
    import torch
    rng = np.random.default_rng(22)
    a = rng.uniform(-1, 1, s.n_cells)
    b = rng.uniform(-1, 1, s.n_cells)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    for _ in range(2):
        assert gpu.dot(da, db) == O.dot(a, b)
        assert gpu.norm_factor(s.b, a) == O.norm_factor(A, s.b, a)


@pytest.mark.parametrize("schedule", [0, 6])
def test_cg_bit_exact(gpu, big, schedule):
    s, A = big
    gpu.set_system(s)
    gpu.set_schedule(schedule, schedule, -1, -1, -1, 1)
    try:
        ref = O.solve("CG", A, s.b, s.x0, 0.0, 6)
        for _ in range(2):
            x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=6, hist_cap=16)
            np.testing.assert_array_equal(hist, ref.hist)
            np.testing.assert_array_equal(x, ref.x)
    finally:
        gpu.set_schedule(0, 0, -1, -1, -1, 1)


def test_mixed_cg_bit_exact(gpu, big):
    s, A = big
    gpu.set_system(s)
    rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 6, 0)
    for _ in range(2):
        x, hist, it, rel, st = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 6, 0)
        assert (st, it) == (rst, rit) == (11, 6)
        np.testing.assert_array_equal(hist, rh)
        np.testing.assert_array_equal(x, rx)


def test_multi_rhs_columns_bit_exact(gpu, big):
    """3 columns: reductions of 3 and 6 totals through the level-2 LDS fold."""
    s, A = big
    gpu.set_system(s)
    rng = np.random.default_rng(23)
    B = rng.uniform(-1, 1, (3, s.n_cells))
    X0 = np.zeros_like(B)
    X, it, nf, sts = gpu.solve_multi("BiCG", B, X0, error=0.0, max_steps=4)
    At = O.transpose(A)
    for k in range(3):
        ref = O.solve("BiCG", A, B[k], X0[k], 0.0, 4, At=At)
        assert it[k] == ref.iterations == 4, k
        assert nf[k] == ref.hist[-1], k
        np.testing.assert_array_equal(X[k], ref.x)


def test_multi_rhs_global_level2_path():
    """330^3 (35,937,000 rows, 35,069 tiles: tree [35069, 548, 9, 1]) with 8
    columns: BiCG's update kernel reduces 16 totals, whose 9 level-2 sums do
    not fit the LDS fold, so the top collector takes the global-partials path.
    Each column must equal its single-RHS solve, whose reductions use the LDS
    path with the same tree, bit for bit."""
    import torch
    from cufoam_amd.gis import Gis
    print("[case] laplacian 330^3 seed=24, 8 columns", flush=True)
    s = mesh.laplacian(330, seed=24, shift=1e-3)
    tiles = -(-s.n_cells // 1024)
    assert (-(-tiles // 64) + 63) // 64 == 9
    with Gis(0) as g:
        g.set_system(s)
        B = torch.from_numpy(np.random.default_rng(25).uniform(-1, 1, (8, s.n_cells))).cuda()
        X, it, nf, sts = g.solve_multi("BiCG", B, torch.zeros_like(B), error=0.0, max_steps=3)
        for k in (0, 5, 7):
            x1, h1, st1 = g.solve("BiCG", B[k].contiguous(), torch.zeros_like(B[k]), error=0.0,
                                  max_steps=3)
            assert it[k] == len(h1) - 1 == 3, k
            assert nf[k] == h1[-1], k
            assert torch.equal(X[k], x1), k


def test_sync_failure_is_reported_and_recovers(gpu):
    """The collectors' bounded wait, forced from the host (gis_set_slot_fault:
    every collector sees its slots unwritten and gives up after 4 polls; the
    kernels still exit normally): dot, NormFactor, single, mixed, multi-RHS and
    batched solves stop with GIS_ERR_SYNC, and after the hook is cleared the refilled slots give
    bit-exact results again."""
    from cufoam_amd.gis import GisError
    s = mesh.laplacian(48, seed=26, shift=1e-2)   # 110,592 rows: 108 tiles, 2 groups
    A = O.ldu2csr(s)
    gpu.set_system(s)
    B = np.random.default_rng(27).uniform(-1, 1, (2, s.n_cells))
    small = [mesh.laplacian(12, seed=28, shift=1e-3), mesh.laplacian(3, seed=29, shift=1e-3)]
    import torch
    b_d = torch.as_tensor(s.b, device="cuda:0")
    gpu.set_slot_fault(4)
    try:
        for call in (lambda: gpu.dot(b_d, b_d), lambda: gpu.norm_factor(b_d, b_d),
                     lambda: gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=5),
                     lambda: gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 5, 0),
                     lambda: gpu.solve_multi("BiCG", B, None, error=0.0, max_steps=5)):
            with pytest.raises(GisError) as e:
                call()
            assert e.value.status == 13
        gpu.batch_set_systems(small)
        bb = np.concatenate([t.b for t in small])
        with pytest.raises(GisError) as e:   # the 2-tile system fails, the 1-tile one cannot
            gpu.batch_solve(bb, None, 1e-9, 500, use_graph=True)
        assert e.value.status == 13
    finally:
        gpu.set_slot_fault(0)
    x, it, nf, st = gpu.batch_solve(bb, None, 1e-9, 500, use_graph=True)
    off = 0
    for k, t in enumerate(small):
        ref = O.solve("CG", O.ldu2csr(t), t.b, t.x0, 1e-9, 500)
        assert it[k] == ref.iterations and nf[k] == ref.hist[-1]
        np.testing.assert_array_equal(x[off:off + t.n_cells], ref.x)
        off += t.n_cells
    gpu.set_system(s)
    assert gpu.dot(b_d, b_d) == O.dot(s.b, s.b)
    assert gpu.norm_factor(b_d, b_d) == O.norm_factor(A, s.b, s.b)
    ref = O.solve("CG", A, s.b, s.x0, 0.0, 5)
    x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=5)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
    rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 5, 0)
    x, hist, it, rel, st = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 5, 0)
    assert (st, it) == (rst, rit)
    np.testing.assert_array_equal(x, rx)
    X, its, nfs, sts = gpu.solve_multi("BiCG", B, None, error=0.0, max_steps=5)
    At = O.transpose(A)
    for k in range(2):
        r1 = O.solve("BiCG", A, B[k], np.zeros(s.n_cells), 0.0, 5, At=At)
        assert nfs[k] == r1.hist[-1]
        np.testing.assert_array_equal(X[k], r1.x)
