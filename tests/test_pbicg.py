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

OpenFOAM's DILU (gis_oracle_amg.c or_dilu_rd/or_dilu_apply) and the LDU form of
the preconditioned BiCG (or_pbicg) are pinned by: M equals A on A's pattern (the
defining property), M z = r and M^T z = r against dense solves, DILU = DIC bit for
bit on symmetric matrices, exact LU on a chain (one-iteration convergence), the
LDU diagonal form = the CSR one bit for bit, PBiCG-DILU = PCG-DIC on symmetric
matrices bit for bit.
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


# ---- DILU (OpenFOAM's DILUPreconditioner, restated in gis_oracle_amg.c) ----------
def _tri(s, lower):
    """Strict lower triangle (lower=True: A(u, l) = lower[f]) or strict upper
    triangle (A(l, u) = upper[f]) of an LduSystem, dense."""
    T = np.zeros((s.n_cells, s.n_cells))
    if lower:
        T[s.upper_addr, s.lower_addr] = s.lower
    else:
        T[s.lower_addr, s.upper_addr] = s.upper
    return T


@pytest.mark.parametrize("n,seed", [(5, 0), (7, 1)])
def test_dilu_matches_a_on_its_pattern_and_inverts_m_and_mt(n, seed):
    """M = (D* + L) D*^-1 (D* + U) with D* = 1 / rD equals A on A's pattern (the
    diagonal, and both coefficients of every face): the defining property of
    DILU.  Its sweeps give z = M^-1 r and, transposed, z = M^-T r (dense solves)."""
    s = mesh.laplacian(n, seed=seed, shift=0.1, symmetric=False)
    rd = O.dilu_rd(s)
    Ds = np.diag(1.0 / rd)
    Lo, Up = _tri(s, True), _tri(s, False)
    M = (Ds + Lo) @ np.diag(rd) @ (Ds + Up)
    A = mesh.ldu_to_dense(s)
    assert not np.allclose(A, A.T)
    np.testing.assert_allclose(np.diag(M), np.diag(A), rtol=1e-12)
    np.testing.assert_allclose(M[s.upper_addr, s.lower_addr], s.lower, rtol=1e-12)
    np.testing.assert_allclose(M[s.lower_addr, s.upper_addr], s.upper, rtol=1e-12)
    r = np.random.default_rng(seed).standard_normal(s.n_cells)
    z = O.dilu_apply(s, rd, r)
    np.testing.assert_allclose(M @ z, r, rtol=1e-10, atol=1e-12 * np.abs(r).max())
    zt = O.dilu_apply(s, rd, r, transpose=True)
    np.testing.assert_allclose(M.T @ zt, r, rtol=1e-10, atol=1e-12 * np.abs(r).max())
    assert not np.allclose(z, zt)


def test_dilu_on_a_symmetric_matrix_is_dic_bit_for_bit():
    s = mesh.laplacian(9, seed=2, shift=1e-3)
    assert np.array_equal(s.lower, s.upper)
    rd = O.dilu_rd(s)
    np.testing.assert_array_equal(rd, O.dic_rd(s))
    r = np.random.default_rng(3).standard_normal(s.n_cells)
    z = O.dic_apply(s, rd, r)
    np.testing.assert_array_equal(O.dilu_apply(s, rd, r), z)
    np.testing.assert_array_equal(O.dilu_apply(s, rd, r, transpose=True), z)


def test_dilu_is_exact_lu_on_a_chain():
    """A 1-D chain has no fill-in, so DILU is the exact LU factorisation and
    preconditioned BiCG converges in one iteration."""
    n = 50
    rng = np.random.default_rng(9)
    la = np.arange(n - 1, dtype=np.int32)
    ua = la + 1
    up = -rng.uniform(0.5, 1.5, n - 1)
    lo = -rng.uniform(0.5, 1.5, n - 1)
    d = np.full(n, 0.3)
    np.add.at(d, la, -up)
    np.add.at(d, ua, -lo)
    s = mesh.LduSystem(n, la, ua, d, lo, up, rng.uniform(-1, 1, n), np.zeros(n))
    A = mesh.ldu_to_dense(s)
    r = rng.standard_normal(n)
    rd = O.dilu_rd(s)
    np.testing.assert_allclose(O.dilu_apply(s, rd, r), np.linalg.solve(A, r), rtol=1e-9)
    np.testing.assert_allclose(O.dilu_apply(s, rd, r, transpose=True),
                               np.linalg.solve(A.T, r), rtol=1e-9)
    res = O.pbicg_ldu("DILU", s, s.b, s.x0, 1e-12 * np.linalg.norm(s.b), 10)
    assert res.status == 0 and res.iterations == 1


@pytest.mark.parametrize("n,seed", [(6, 0), (11, 1)])
def test_ldu_pbicg_diagonal_is_the_csr_restatement(n, seed):
    """The LDU form (or_pbicg) and the CSR form (pbicg_diag) of diagonally
    preconditioned BiCG are the same iteration, bit for bit."""
    s = mesh.laplacian(n, seed=seed, shift=0.05, symmetric=False)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 2.0, s.n_cells)
    a = O.pbicg_ldu("diagonal", s, s.b, s.x0, 0.0, 20)
    b = O.pbicg("diagonal", O.ldu2csr(s), s.b, s.x0, 0.0, 20)
    np.testing.assert_array_equal(a.hist, b.hist)
    np.testing.assert_array_equal(a.x, b.x)


def test_pbicg_dilu_on_a_symmetric_matrix_is_pcg_dic():
    s = mesh.laplacian(10, seed=5, shift=1e-3)
    for error, steps in ((0.0, 25), (1e-9 * np.linalg.norm(s.b), 2000)):
        a = O.pbicg_ldu("DILU", s, s.b, s.x0, error, steps)
        b = O.pcg("DIC", s, s.b, s.x0, error, steps)
        assert a.hist_len == b.hist_len
        np.testing.assert_array_equal(a.hist, b.hist)
        np.testing.assert_array_equal(a.x, b.x)


def test_pbicg_dilu_beats_diagonal_on_a_nonsymmetric_matrix():
    s = mesh.laplacian(16, seed=6, shift=1e-3, symmetric=False)
    err = 1e-8 * np.linalg.norm(s.b)
    dl = O.pbicg_ldu("DILU", s, s.b, s.x0, err, 5000)
    dg = O.pbicg_ldu("diagonal", s, s.b, s.x0, err, 5000)
    assert dl.status == dg.status == 0
    assert dl.iterations < 0.7 * dg.iterations
    assert np.linalg.norm(s.b - mesh.ldu_matvec(s, dl.x)) <= 1.01 * err
