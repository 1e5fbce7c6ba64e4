"""The OpenMP restatements of CG and mixed CG (oracle/gis_oracle_par.c) give
the serial oracle's bits: same histories, same solutions, same iteration
counts, on sizes that leave ragged leaves, tiles and tile groups."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O


@pytest.mark.parametrize("n,seed,shift", [(1, 0, 1e-3), (3, 1, 1e-3), (10, 2, 1e-2),
                                          (11, 3, 1e-3), (17, 4, 1e-3), (41, 5, 1e-4)])
def test_cg_par_equals_serial_cg(n, seed, shift):
    s = mesh.laplacian(n, seed=seed, shift=shift)
    A = O.ldu2csr(s)
    x0 = np.random.default_rng(seed).uniform(-0.5, 0.5, s.n_cells) if seed % 2 else s.x0
    err = 1e-9 * float(np.linalg.norm(s.b))
    ref = O.solve("CG", A, s.b, x0, err, 3000)
    got = O.cg_par(A, s.b, x0, err, 3000)
    assert got.hist_len == ref.hist_len
    np.testing.assert_array_equal(got.hist, ref.hist)
    np.testing.assert_array_equal(got.x, ref.x)
    fixed = O.cg_par(A, s.b, x0, 0.0, 7)
    assert fixed.hist_len == 8
    np.testing.assert_array_equal(fixed.x, O.solve("CG", A, s.b, x0, 0.0, 7).x)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,seed", [(9, 6), (33, 7)])
def test_mixed_cg_par_equals_serial(n, seed, mode):
    s = mesh.laplacian(n, seed=seed, shift=1e-4)
    A = O.ldu2csr(s)
    for inner, cap in ((1e-2, 100000), (1e-4, 100000), (1e-2, 17)):
        rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, inner, cap, mode)
        gx, gh, git, gst = O.mixed_cg_par(A, s.b, s.x0, 1e-10, inner, cap, mode)
        assert (git, gst) == (rit, rst)
        np.testing.assert_array_equal(gh, rh)
        np.testing.assert_array_equal(gx, rx)


def test_residual_norm_par_matches_the_oracle_residual():
    s = mesh.laplacian(23, seed=8, shift=1e-3, symmetric=False)
    A = O.ldu2csr(s)
    x = np.random.default_rng(9).uniform(-1, 1, s.n_cells)
    r = O.residual(A, s.b, x)
    assert O.residual_norm_par(A, s.b, x) == np.sqrt(O.dot(r, r))
