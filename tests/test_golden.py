"""Golden solver fixtures (tests/golden/{hex,random}*.npz, written by
tests/golden/make_golden.py):
the CPU oracle must reproduce them bit for bit on any host (no GPU), and the
HIP path through the C ABI must reproduce them bit for bit on the GPU."""
import glob
import os

import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "hex*.npz")) +
                  glob.glob(os.path.join(HERE, "golden", "random*.npz")))
METHODS = ["CG", "BiCG", "Jacobi", "MR", "MRJacobi", "JMR", "MRS", "Richardson"]
STEPS = 20


def _load(path):
    z = np.load(path)   # allow_pickle=False: data only
    s = mesh.LduSystem(int(z["n_cells"]), z["lower_addr"], z["upper_addr"], z["diag"],
                       z["lower"], z["upper"], z["b"], z["x0"])
    return z, s


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    z, s = _load(path)
    A = O.ldu2csr(s)
    assert np.array_equal(A.rowptr, z["rowptr"]) and np.array_equal(A.col, z["col"])
    assert np.array_equal(_bits(A.val), _bits(z["val"]))
    for m in METHODS:
        r = O.solve(m, A, s.b, z["x0"], 0.0, STEPS)
        assert np.array_equal(_bits(r.hist), _bits(z[f"{m}_hist"])), m
        assert np.array_equal(_bits(r.x), _bits(z[f"{m}_x"])), m
    if "mixed0_x" in z:
        s2 = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, z["mixed0_diag"], s.lower,
                            s.upper, s.b, np.zeros(s.n_cells))
        A2 = O.ldu2csr(s2)
        for mode in (0, 1):
            x, h, it, st = O.mixed_cg(A2, s2.b, s2.x0, 1e-10, 1e-2, 5000, mode)
            assert it == int(z[f"mixed{mode}_iters"])
            assert np.array_equal(_bits(h), _bits(z[f"mixed{mode}_hist"]))
            assert np.array_equal(_bits(x), _bits(z[f"mixed{mode}_x"]))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_golden(gpu, path):
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    z, s = _load(path)
    gpu.set_system(s)
    rp, col, val = gpu.get_csr()
    assert np.array_equal(rp, z["rowptr"]) and np.array_equal(col, z["col"])
    assert np.array_equal(_bits(val), _bits(z["val"]))
    for m in METHODS:
        x, hist, st = gpu.solve(m, s.b, z["x0"], error=0.0, max_steps=STEPS, hist_cap=STEPS + 1)
        assert np.array_equal(_bits(hist), _bits(z[f"{m}_hist"])), m
        assert np.array_equal(_bits(x), _bits(z[f"{m}_x"])), m
    if "mixed0_x" in z:
        gpu.set_values(z["mixed0_diag"], s.lower, s.upper)
        for mode in (0, 1):
            x, h, it, rel, st = gpu.solve_mixed(s.b, np.zeros(s.n_cells), 1e-10, 1e-2, 5000,
                                                mode)
            assert it == int(z[f"mixed{mode}_iters"])
            assert np.array_equal(_bits(h), _bits(z[f"mixed{mode}_hist"]))
            assert np.array_equal(_bits(x), _bits(z[f"mixed{mode}_x"]))
