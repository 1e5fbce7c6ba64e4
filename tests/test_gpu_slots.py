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
