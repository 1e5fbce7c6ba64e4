"""Pinning the oracle's preconditioned CG (none / diagonal / GAMG) on the CPU.

No counterpart in the reference (its roadmap lists preconditioners as future
work, README.md:71-73), so the restatement is pinned by what the mathematics
fixes:
  * PCG with no preconditioner is gis CG bit for bit;
  * diagonal PCG follows a textbook PCG (numpy) and, for a constant diagonal,
    CG's iterates;
  * the pairwise pass on small graphs gives the hand-derived aggregates;
  * every coarse level equals P^T A P for the 0/1 aggregation P, aggregates
    partition the cells and are connected, coarse faces are distinct and in
    upper-triangular order; the last level's dense inverse inverts it;
  * the V-cycle is a symmetric positive definite operator (so PCG applies),
    and GAMG-PCG solves the system in far fewer iterations than CG;
  * the stopping rule with relTol / minIter stops where predicted.
"""
import numpy as np
import pytest
import scipy.sparse.csgraph as csgraph
import scipy.sparse as sp

from cufoam_amd import mesh
from oracle import oracle as O


def _dense_level(lv):
    n = lv["n"]
    D = np.diag(lv["diag"].copy())
    np.add.at(D, (lv["la"], lv["ua"]), lv["upper"])
    np.add.at(D, (lv["ua"], lv["la"]), lv["lower"])
    return D


@pytest.mark.parametrize("symmetric", [True, False])
def test_pcg_none_is_cg_bitwise(symmetric):
    s = mesh.laplacian(9, seed=1, shift=0.05, symmetric=symmetric)
    A = O.ldu2csr(s)
    x0 = np.random.default_rng(2).uniform(-1, 1, s.n_cells)
    ref = O.solve("CG", A, s.b, x0, 1e-9, 300)
    got = O.pcg("none", s, s.b, x0, 1e-9, 300)
    np.testing.assert_array_equal(got.hist, ref.hist)
    np.testing.assert_array_equal(got.x, ref.x)


def _textbook_pcg(A, b, x, Minv, k):
    r = b - A @ x
    z = Minv * r
    p = z.copy()
    rho = r @ z
    hist = [np.linalg.norm(r)]
    for _ in range(k):
        q = A @ p
        a = rho / (p @ q)
        x = x + a * p
        r = r - a * q
        hist.append(np.linalg.norm(r))
        z = Minv * r
        rn = r @ z
        p = z + (rn / rho) * p
        rho = rn
    return x, np.array(hist)


def test_diagonal_pcg_follows_the_textbook_recurrence():
    s = mesh.laplacian(8, seed=3, shift=0.3)
    s.diag = s.diag * np.random.default_rng(4).uniform(1.0, 3.0, s.n_cells)   # uneven diagonal
    A = O.ldu2csr(s).scipy()
    got = O.pcg("diagonal", s, s.b, s.x0, 0.0, 40)
    x, hist = _textbook_pcg(A, s.b, s.x0.copy(), 1.0 / s.diag, 40)
    np.testing.assert_allclose(got.hist, hist, rtol=1e-8)
    np.testing.assert_allclose(got.x, x, rtol=1e-8, atol=1e-12)
    cg = O.pcg("none", s, s.b, s.x0, 0.0, 40)
    assert not np.allclose(got.hist, cg.hist)   # the preconditioner acts


def test_diagonal_pcg_on_a_constant_diagonal_is_cg():
    s = mesh.laplacian(6, seed=5, shift=0.0)
    s.diag = np.full(s.n_cells, 8.0)
    got = O.pcg("diagonal", s, s.b, s.x0, 0.0, 20)
    cg = O.pcg("none", s, s.b, s.x0, 0.0, 20)
    np.testing.assert_allclose(got.hist, cg.hist, rtol=1e-12)


def test_pair_pass_known_answers():
    # chain 0-1-2-3-4-5-6-7, weights 1,5,1,1,5,1,1: cells pair in visiting order
    la, ua = np.arange(7), np.arange(1, 8)
    nc, agg = O.pair_pass(8, la, ua, [1, 5, 1, 1, 5, 1, 1])
    assert nc == 4 and agg.tolist() == [0, 0, 1, 1, 2, 2, 3, 3]
    # star 0-{1,2,3} with weights 1,3,2: 0 pairs with its heaviest neighbour 2,
    # then 1 and 3 (no unpaired neighbour) join the aggregate of 0
    nc, agg = O.pair_pass(4, [0, 0, 0], [1, 2, 3], [1, 3, 2])
    assert nc == 1 and agg.tolist() == [0, 0, 0, 0]
    # 0-1-2 chain with a pendant 3 on 1: 2 and 3 find 1 taken and join its aggregate
    nc, agg = O.pair_pass(4, [0, 1, 1], [1, 2, 3], [1, 2, 2])
    assert nc == 1 and agg.tolist() == [0, 0, 0, 0]
    # ties among unpaired neighbours: the first in row order wins (0 pairs with 1,
    # so 2 pairs with 3; choosing 2 would leave one aggregate)
    nc, agg = O.pair_pass(4, [0, 0, 2], [1, 2, 3], [2, 2, 1])
    assert nc == 2 and agg.tolist() == [0, 0, 1, 1]
    # ties when joining: faces where the cell is neighbour come first in its row
    # (cell 2 sees 1 via face (1,2) before 3 via face (2,3), both weight 2)
    nc, agg = O.pair_pass(5, [0, 1, 1, 2], [3, 2, 4, 3], [5, 2, 5, 2])
    assert nc == 2 and agg.tolist() == [0, 1, 1, 0, 1]
    # isolated cells stay alone
    nc, agg = O.pair_pass(3, [], [], [])
    assert nc == 3 and agg.tolist() == [0, 1, 2]


@pytest.mark.parametrize("n,symmetric", [(8, True), (12, False), (16, True)])
def test_hierarchy_is_galerkin_and_well_formed(n, symmetric):
    s = mesh.laplacian(n, seed=n, shift=0.01, symmetric=symmetric)
    a = O.Amg(s)
    assert a.n_levels >= 2
    prev = a.level(0)
    np.testing.assert_array_equal(_dense_level(prev), mesh.ldu_to_dense(s))
    for l in range(1, a.n_levels):
        cur = a.level(l)
        agg = prev["agg"]
        assert cur["n"] == prev["nc"] and set(agg.tolist()) == set(range(cur["n"]))
        assert cur["n"] <= prev["n"] * 9 // 10
        P = np.zeros((prev["n"], cur["n"]))
        P[np.arange(prev["n"]), agg] = 1.0
        np.testing.assert_allclose(_dense_level(cur), P.T @ _dense_level(prev) @ P,
                                   rtol=1e-13, atol=1e-12)
        key = cur["la"].astype(np.int64) * cur["n"] + cur["ua"]
        assert np.all(cur["la"] < cur["ua"]) and np.all(np.diff(key) > 0)
        # each aggregate is connected in the fine face graph
        G = sp.coo_matrix((np.ones(prev["m"]), (prev["la"], prev["ua"])),
                          shape=(prev["n"], prev["n"])).tocsr()
        for c in range(0, cur["n"], max(1, cur["n"] // 40)):
            idx = np.nonzero(agg == c)[0]
            k, _ = csgraph.connected_components(G[idx][:, idx], directed=False)
            assert k == 1
        prev = cur
    assert prev["n"] <= 64
    Ainv = a.dense_inverse()
    np.testing.assert_allclose(Ainv @ _dense_level(prev), np.eye(prev["n"]), atol=1e-10)


def test_vcycle_is_symmetric_positive_definite():
    s = mesh.laplacian(6, seed=7, shift=0.01)
    a = O.Amg(s)
    assert a.n_levels >= 2
    n = s.n_cells
    Mi = np.stack([a.vcycle(e) for e in np.eye(n)], axis=1)
    np.testing.assert_allclose(Mi, Mi.T, rtol=1e-10, atol=1e-13)
    assert np.linalg.eigvalsh((Mi + Mi.T) / 2).min() > 0
    # linear: V(a r1 + r2) = a V(r1) + V(r2)
    r1, r2 = np.random.default_rng(1).standard_normal((2, n))
    np.testing.assert_allclose(a.vcycle(2.5 * r1 + r2), 2.5 * a.vcycle(r1) + a.vcycle(r2),
                               rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("n", [16, 24])
def test_gamg_pcg_converges_fast_to_the_solution(n):
    s = mesh.laplacian(n, seed=8, shift=1e-4)
    err = 1e-10 * np.linalg.norm(s.b)
    cg = O.solve("CG", O.ldu2csr(s), s.b, s.x0, err, 10000)
    g = O.pcg("GAMG", s, s.b, s.x0, err, 10000)
    assert g.hist[-1] <= err and g.iterations * 4 < cg.iterations
    A = O.ldu2csr(s).scipy()
    assert np.linalg.norm(s.b - A @ g.x) <= 2 * err


def test_gamg_single_level_and_stagnation():
    s = mesh.laplacian(4, seed=9, shift=0.1)          # 64 cells: one level, exact inverse
    a = O.Amg(s)
    assert a.n_levels == 1
    g = O.pcg("GAMG", s, s.b, s.x0, 1e-9 * np.linalg.norm(s.b), 50)
    assert g.iterations == 1
    n = 300                                            # no faces: nothing to agglomerate
    d = mesh.LduSystem(n, np.zeros(0, np.int32), np.zeros(0, np.int32),
                       np.random.default_rng(3).uniform(1, 2, n), np.zeros(0), np.zeros(0),
                       np.random.default_rng(4).uniform(-1, 1, n), np.zeros(n))
    assert O.Amg(d).n_levels == 1
    g = O.pcg("GAMG", d, d.b, d.x0, 1e-12, 50)
    assert g.iterations == 1


def test_pcg_rel_tol_and_min_steps():
    s = mesh.laplacian(10, seed=10, shift=0.01)
    long = O.pcg("GAMG", s, s.b, s.x0, 0.0, 30)
    h = long.hist
    thr = 1e-3 * h[0]
    k = next(i for i in range(len(h)) if not h[i] > thr)
    r = O.pcg("GAMG", s, s.b, s.x0, 0.0, 100, rel_tol=1e-3)
    assert r.iterations == k
    np.testing.assert_array_equal(r.hist, h[:k + 1])
    r = O.pcg("GAMG", s, s.b, s.x0, 1e30, 100, min_steps=4)
    assert r.iterations == 4
    np.testing.assert_array_equal(r.hist, h[:5])


def test_fp32_cycle_tracks_the_fp64_cycle_and_converges():
    s = mesh.laplacian(20, seed=11, shift=1e-3)
    a = O.Amg(s)
    r = np.random.default_rng(2).standard_normal(s.n_cells)
    z64, z32 = a.vcycle(r), a.vcycle32(r)
    assert 1e-12 < np.linalg.norm(z32 - z64) / np.linalg.norm(z64) < 1e-5
    # it is the FP64 cycle applied in float: the same operator up to float rounding
    z32b = a.vcycle32(2.0 * r)
    np.testing.assert_array_equal(z32b, 2.0 * z32)   # scaling by 2 is exact in float
    err = 1e-10 * np.linalg.norm(s.b)
    g64 = O.pcg("GAMG", s, s.b, s.x0, err, 1000)
    g32 = O.pcg("GAMG32", s, s.b, s.x0, err, 1000)
    assert g32.hist[-1] <= err and g32.iterations <= g64.iterations + 3
    A = O.ldu2csr(s).scipy()
    assert np.linalg.norm(s.b - A @ g32.x) <= 2 * err


# ---- DIC (OpenFOAM's DICPreconditioner, restated in gis_oracle_amg.c) ----------------
def _lower_dense(s):
    """Strict lower triangle of A in dense form (A(u, l) = lower, here = upper)."""
    Lm = np.zeros((s.n_cells, s.n_cells))
    Lm[s.upper_addr, s.lower_addr] = s.upper
    return Lm


@pytest.mark.parametrize("n,shift", [(5, 0.1), (7, 1e-3)])
def test_dic_is_incomplete_cholesky_on_the_pattern(n, shift):
    """M = (D* + L) D*^-1 (D* + L^T) with D* = 1 / rD: M equals A on A's pattern
    (diagonal and every face), the defining property of IC(0); and z = M^-1 r
    for the oracle's sweeps, against a dense solve."""
    from cufoam_amd import mesh
    s = mesh.laplacian(n, seed=3, shift=shift)
    rd = O.dic_rd(s)
    Ds = np.diag(1.0 / rd)
    Lm = _lower_dense(s)
    M = (Ds + Lm) @ np.diag(rd) @ (Ds + Lm.T)
    A = mesh.ldu_to_dense(s)
    np.testing.assert_allclose(np.diag(M), np.diag(A), rtol=1e-12)
    np.testing.assert_allclose(M[s.upper_addr, s.lower_addr], A[s.upper_addr, s.lower_addr],
                               rtol=1e-12)
    r = np.random.default_rng(5).standard_normal(s.n_cells)
    z = O.dic_apply(s, rd, r)
    np.testing.assert_allclose(M @ z, r, rtol=1e-10, atol=1e-12 * np.abs(r).max())


def test_dic_is_exact_cholesky_on_a_chain_and_diagonal_without_faces():
    from cufoam_amd import mesh
    n = 40   # a 1-D chain: no fill-in, so IC(0) is the exact factorisation
    la = np.arange(n - 1, dtype=np.int32)
    ua = la + 1
    up = -np.random.default_rng(6).uniform(0.5, 1.5, n - 1)
    d = np.full(n, 0.05)
    np.add.at(d, la, -up)
    np.add.at(d, ua, -up)
    s = mesh.LduSystem(n, la, ua, d, up.copy(), up, np.ones(n), np.zeros(n))
    r = np.random.default_rng(7).standard_normal(n)
    z = O.dic_apply(s, O.dic_rd(s), r)
    np.testing.assert_allclose(z, np.linalg.solve(mesh.ldu_to_dense(s), r), rtol=1e-9)
    e = mesh.LduSystem(5, np.zeros(0, np.int32), np.zeros(0, np.int32), np.arange(1.0, 6.0),
                       np.zeros(0), np.zeros(0), np.ones(5), np.zeros(5))
    assert np.array_equal(O.dic_rd(e), 1.0 / np.arange(1.0, 6.0))


def test_dic_pcg_converges_faster_than_cg():
    from cufoam_amd import mesh
    s = mesh.laplacian(16, seed=4, shift=1e-3)
    err = 1e-8 * np.linalg.norm(s.b)
    plain = O.pcg("none", s, s.b, s.x0, err, 5000)
    dic = O.pcg("DIC", s, s.b, s.x0, err, 5000)
    assert dic.iterations < 0.7 * plain.iterations
    A = mesh.ldu_to_dense(s)
    assert np.linalg.norm(s.b - A @ dic.x) <= 1.01 * err
