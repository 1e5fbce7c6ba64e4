"""The oracle's BiCG preconditioned by the diagonal (oracle/gis_oracle.c
pbicg_diag; no counterpart in the reference, whose BiCG, SolverSingleGPU.h:568-675,
takes no preconditioner), pinned four ways:

  - a unit diagonal makes z = 1.0 * r = r exactly, so it must be the oracle's
    BiCG bit for bit (a dropped or misplaced D^-1 changes nothing there, which
    the next two checks cover);
  - on a symmetric matrix r* = r and A^T = A as stored, so its iterates must be
    those of PCG with the diagonal preconditioner (gis_oracle_amg.c or_pcg) bit
    for bit;
  - on a nonsymmetric matrix with a non-unit diagonal it must follow the
    textbook preconditioned BiCG (Barrett et al., "Templates for the Solution of
    Linear Systems", SIAM 1994, fig. 2.8: z = M^-1 r, z~ = M^-T r~, rho = z.r~,
    p = z + beta p, p~ = z~ + beta p~, alpha = rho / p~.Ap) written here in
    dense numpy, to rounding; z~ taken from r instead of r~, or D^-1 left out of
    rho, fails it;
  - finite termination: N iterations solve an N-cell system (exact arithmetic).
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O


def _unit_diagonal(s):
    """Rows scaled by 1 / a_ii: lower[f] sits in row upperAddr[f], upper[f] in row
    lowerAddr[f]."""
    d = s.diag
    return mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, np.ones(s.n_cells),
                          s.lower / d[s.upper_addr], s.upper / d[s.lower_addr], s.b / d,
                          np.zeros(s.n_cells))


@pytest.mark.parametrize("n,seed", [(4, 0), (7, 1)])
def test_unit_diagonal_is_bicg_bit_for_bit(n, seed):
    s = _unit_diagonal(mesh.laplacian(n, seed=seed, shift=0.1, symmetric=False))
    assert np.all(s.diag == 1.0)
    A = O.ldu2csr(s)
    x0 = np.random.default_rng(seed).uniform(-1, 1, s.n_cells)
    ref = O.solve("BiCG", A, s.b, x0, 0.0, 15)
    got = O.pbicg("diagonal", A, s.b, x0, 0.0, 15)
    assert got.status == ref.status == 0
    np.testing.assert_array_equal(got.hist, ref.hist)
    np.testing.assert_array_equal(got.x, ref.x)


@pytest.mark.parametrize("n,seed", [(3, 0), (8, 2), (12, 3)])
def test_symmetric_is_pcg_diagonal_bit_for_bit(n, seed):
    s = mesh.laplacian(n, seed=seed, shift=0.05)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 2.0, s.n_cells)
    A = O.ldu2csr(s)
    for error, steps in ((0.0, 20), (1e-9 * np.linalg.norm(s.b), 500)):
        ref = O.pcg("diagonal", s, s.b, s.x0, error, steps)
        got = O.pbicg("diagonal", A, s.b, s.x0, error, steps)
        assert got.hist_len == ref.hist_len
        np.testing.assert_array_equal(got.hist, ref.hist)
        np.testing.assert_array_equal(got.x, ref.x)


def _textbook(Ad, b, x0, k):
    minv = 1.0 / np.diag(Ad)
    x = x0.copy()
    r = b - Ad @ x
    rt = r.copy()
    hist = [np.linalg.norm(r)]
    p = pt = None
    rho_prev = 0.0
    for i in range(k):
        z, zt = minv * r, minv * rt
        rho = z @ rt
        if i == 0:
            p, pt = z.copy(), zt.copy()
        else:
            beta = rho / rho_prev
            p, pt = z + beta * p, zt + beta * pt
        q, qt = Ad @ p, Ad.T @ pt
        alpha = rho / (pt @ q)
        x, r, rt = x + alpha * p, r - alpha * q, rt - alpha * qt
        rho_prev = rho
        hist.append(np.linalg.norm(r))
    return x, np.array(hist)


@pytest.mark.parametrize("n_cells,n_faces,seed", [(60, 150, 1), (200, 700, 2)])
def test_nonsymmetric_follows_the_textbook_iteration(n_cells, n_faces, seed):
    s = mesh.random_ldu(n_cells, n_faces, seed=seed, symmetric=False, shift=0.3)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 3.0, n_cells)
    Ad = mesh.ldu_to_dense(s)
    assert not np.allclose(Ad, Ad.T) and np.ptp(np.diag(Ad)) > 1.0
    k = 8
    got = O.pbicg("diagonal", O.ldu2csr(s), s.b, s.x0, 0.0, k)
    x, hist = _textbook(Ad, s.b, s.x0, k)
    assert got.hist_len == k + 1
    np.testing.assert_allclose(got.hist, hist, rtol=1e-9)
    np.testing.assert_allclose(got.x, x, rtol=1e-9, atol=1e-12)
    plain = O.solve("BiCG", O.ldu2csr(s), s.b, s.x0, 0.0, k)
    assert not np.allclose(plain.hist, hist, rtol=1e-6)   # the preconditioner matters here


def test_finite_termination():
    s = mesh.random_ldu(12, 30, seed=4, symmetric=False, shift=0.2)
    s.diag = s.diag * np.random.default_rng(4).uniform(1.0, 2.0, 12)
    A = O.ldu2csr(s)
    bn = np.linalg.norm(s.b)
    got = O.pbicg("diagonal", A, s.b, s.x0, 1e-11 * bn, 12, int_bound=0.0)
    assert got.status == 0 and got.iterations <= 12
    Ad = mesh.ldu_to_dense(s)
    assert np.linalg.norm(s.b - Ad @ got.x) <= 1e-10 * bn


def test_other_preconditioners_are_rejected():
    s = mesh.laplacian(3, seed=0)
    with pytest.raises(ValueError):
        O.pbicg("GAMG", O.ldu2csr(s), s.b, s.x0, 0.0, 3)
