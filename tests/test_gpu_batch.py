"""GPU: batched CG over independent systems of uneven size.

Every system of a batch must reproduce its own single-system CG bit for bit
(solution, iteration count, final residual norm), with and without the
captured hipGraph, across repeated graph replays with other work in between."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _systems(edges, seed0=0, shift=1e-3):
    return [mesh.laplacian(int(n), seed=seed0 + i, shift=shift) for i, n in enumerate(edges)]


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("schedule", [None, 6, 0])
@pytest.mark.parametrize("direction", [0, 1])
def test_batch_bit_exact_vs_oracle(gpu, use_graph, schedule, direction):
    """direction 0: the direction formed at the gathers (k_bcg1); 1: the split
    form (k_bdir + k_bcg1s), which the pattern picks for large stacks."""
    edges = [2, 5, 3, 12, 1, 9, 7]
    systems = _systems(edges, 10)
    print(f"[case] batch edges={edges} graph={use_graph} schedule={schedule} "
          f"direction={direction}", flush=True)
    gpu.batch_set_systems(systems)
    gpu.set_schedule(spmv64=-1 if schedule is None else schedule, fp64_direction=direction)
    b = np.concatenate([s.b for s in systems])
    err = np.array([1e-9 * np.linalg.norm(s.b) for s in systems])
    for rep in range(3):
        x, it, nf, st = gpu.batch_solve(b, None, err, 1000, use_graph=use_graph)
        off = 0
        for k, s in enumerate(systems):
            ref = O.solve("CG", O.ldu2csr(s), s.b, s.x0, err[k], 1000)
            assert it[k] == ref.iterations, (rep, k)
            assert nf[k] == ref.hist[-1], (rep, k)
            np.testing.assert_array_equal(x[off:off + s.n_cells], ref.x)
            off += s.n_cells


def test_batch_graph_replays_with_interleaved_work(gpu):
    import torch
    systems = _systems([6, 11, 4], 20)
    gpu.batch_set_systems(systems)
    b = np.concatenate([s.b for s in systems])
    err = np.full(3, 1e-10)
    x0, it0, nf0, _ = gpu.batch_solve(b, None, err, 2000, use_graph=False)
    junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
    for rep in range(4):
        float(junk.sum())   # other work between replays
        x, it, nf, _ = gpu.batch_solve(b, None, err, 2000, use_graph=True)
        assert np.array_equal(it, it0) and np.array_equal(nf, nf0)
        np.testing.assert_array_equal(x, x0)


def test_batch_64_uneven_systems_equal_single_solves(gpu):
    from cufoam_amd.gis import Gis
    edges = mesh.batch_edges(64, 16, 64, seed=0)
    systems = _systems(edges, 100)
    N = sum(s.n_cells for s in systems)
    print(f"[case] 64 systems, {N} cells in total", flush=True)
    gpu.batch_set_systems(systems)
    b = np.concatenate([s.b for s in systems])
    err = np.array([1e-8 * np.linalg.norm(s.b) for s in systems])
    x, it, nf, st = gpu.batch_solve(b, None, err, 5000, use_graph=True)
    assert np.all(st == 0)
    with Gis(0) as single:
        off = 0
        for k, s in enumerate(systems):
            single.set_system(s)
            xs, hist, _ = single.solve("CG", s.b, s.x0, error=err[k], max_steps=5000,
                                       hist_cap=5001)
            assert it[k] == len(hist) - 1 and nf[k] == hist[-1], k
            np.testing.assert_array_equal(x[off:off + s.n_cells], xs)
            off += s.n_cells
