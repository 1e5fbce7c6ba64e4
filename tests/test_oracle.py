"""Pinning the CPU oracle (no GPU).

The cuFoam reference ships no tests, fixtures or golden vectors, its solvers
need CUDA and OpenFOAM (absent here), so the oracle is pinned by
  * hand-derived known answers (tiny meshes, reduction trees),
  * independent restatements of the same layout (the counting-sort variant
    left commented out in gis/ldu2csr.h:43-95; closed forms for row offsets),
  * the reference's own CPU code compiled from its sources
    (SparseCPU.h transpose / JacobiPrecondition, oracle/_ref),
  * library routines (scipy.sparse CSR matvec, scipy cg / bicg, dense solves),
  * mathematical invariants (BiCG == CG on symmetric A, MRS == MRJacobi,
    recursive residual == true residual, one-step convergence on c*I, ...),
each chosen so that a plausible mistake (swapped lower/upper, wrong sign,
dropped term, transposed operand, wrong summation order) fails one of them.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cufoam_amd import mesh
from oracle import oracle as O


def _dense(A: O.Csr) -> np.ndarray:
    return A.scipy().toarray()


# ---------------------------------------------------------------------------
# ldu2csr (gis/ldu2csr.h:19-39)
# ---------------------------------------------------------------------------
def test_ldu2csr_known_answer_4_cells():
    # faces (0,1) (0,2) (1,3) (2,3); row c = [lower faces of c by face index | diag | upper faces]
    la = np.array([0, 0, 1, 2], np.int32)
    ua = np.array([1, 2, 3, 3], np.int32)
    s = mesh.LduSystem(4, la, ua, np.array([10., 11., 12., 13.]),
                       np.array([-1., -2., -3., -4.]), np.array([-5., -6., -7., -8.]),
                       np.zeros(4), np.zeros(4))
    A = O.ldu2csr(s)
    assert A.rowptr.tolist() == [0, 3, 6, 9, 12]
    assert A.col.tolist() == [0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3]
    assert A.val.tolist() == [10., -5., -6., -1., 11., -7., -2., 12., -8., -3., -4., 13.]


def test_ldu2csr_known_answer_losort_interleaves_owners():
    # faces (0,3) (1,2) (1,3) (2,3): row 3 gathers lower faces 0, 2, 3 (owners 0, 1, 2)
    la = np.array([0, 1, 1, 2], np.int32)
    ua = np.array([3, 2, 3, 3], np.int32)
    s = mesh.LduSystem(4, la, ua, np.array([1., 2., 3., 4.]),
                       np.array([10., 20., 30., 40.]), np.array([50., 60., 70., 80.]),
                       np.zeros(4), np.zeros(4))
    A = O.ldu2csr(s)
    assert A.rowptr.tolist() == [0, 2, 5, 8, 12]
    assert A.col.tolist() == [0, 3, 1, 2, 3, 1, 2, 3, 0, 1, 2, 3]
    assert A.val.tolist() == [1., 50., 2., 60., 70., 20., 3., 80., 10., 30., 40., 4.]


def _counting_variant(s):
    """The earlier counting-sort formulation kept (commented out) in gis/ldu2csr.h:43-95:
    row lengths by counting, lower entries scattered in face order to their neighbour
    row, then the diagonal, then upper entries in face order to their owner row."""
    N = s.n_cells
    counts = np.ones(N + 1, np.int64)
    counts[0] = 0
    np.add.at(counts, s.lower_addr + 1, 1)
    np.add.at(counts, s.upper_addr + 1, 1)
    rp = np.cumsum(counts)
    fill = rp[:-1].copy()
    col = np.zeros(rp[-1], np.int64)
    val = np.zeros(rp[-1])
    for f in range(s.n_faces):
        r = s.upper_addr[f]
        col[fill[r]] = s.lower_addr[f]
        val[fill[r]] = s.lower[f]
        fill[r] += 1
    for c in range(N):
        col[fill[c]] = c
        val[fill[c]] = s.diag[c]
        fill[c] += 1
    for f in range(s.n_faces):
        r = s.lower_addr[f]
        col[fill[r]] = s.upper_addr[f]
        val[fill[r]] = s.upper[f]
        fill[r] += 1
    return rp, col, val


@pytest.mark.parametrize("kind", ["hex4", "hex7", "random", "random_dense"])
def test_ldu2csr_equals_counting_variant_and_dense(kind):
    if kind == "hex4":
        s = mesh.laplacian(4, seed=1, symmetric=False)
    elif kind == "hex7":
        s = mesh.laplacian(7, seed=2, symmetric=False)
    elif kind == "random":
        s = mesh.random_ldu(300, 900, seed=3)
    else:
        s = mesh.random_ldu(40, 500, seed=4)
    A = O.ldu2csr(s)
    rp, col, val = _counting_variant(s)
    assert np.array_equal(A.rowptr, rp) and np.array_equal(A.col, col)
    assert np.array_equal(A.val, val)
    assert np.array_equal(_dense(A), mesh.ldu_to_dense(s))
    # columns strictly ascending in every row for upper-triangular face order
    for i in range(A.n):
        seg = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("n,seed", [(3, 0), (16, 1)])
def test_ldu2csr_row_offset_closed_form(n, seed):
    s = mesh.laplacian(n, seed=seed)
    A = O.ldu2csr(s)
    cells = np.arange(s.n_cells)
    owner_start = np.searchsorted(s.lower_addr, cells, side="left")
    losort_start = np.searchsorted(np.sort(s.upper_addr, kind="stable"), cells, side="left")
    assert np.array_equal(A.rowptr[:-1], cells + losort_start + owner_start)
    assert A.rowptr[-1] == s.nnz


def test_ldu2csr_rejects_invalid_addressing():
    with pytest.raises(ValueError):
        O.ldu2csr_pattern(3, np.array([1, 0], np.int32), np.array([2, 2], np.int32))
    with pytest.raises(ValueError):
        O.ldu2csr_pattern(3, np.array([0, 1], np.int32), np.array([3, 2], np.int32))


def test_ldu_transpose_closed_form_and_reference_transpose():
    s = mesh.random_ldu(200, 700, seed=5)
    A = O.ldu2csr(s)
    At = O.transpose(A)
    swapped = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, s.upper, s.lower,
                             s.b, s.x0)
    B = O.ldu2csr(swapped)
    assert np.array_equal(At.rowptr, B.rowptr) and np.array_equal(At.col, B.col)
    assert np.array_equal(At.val, B.val)
    assert np.array_equal(_dense(At), _dense(A).T)
    Att = O.transpose(At)
    assert np.array_equal(Att.col, A.col) and np.array_equal(Att.val, A.val)
    if O.ref_lib() is None:
        pytest.skip("reference SparseCPU.h driver not built")
    R = O.ref_transpose(A)
    assert np.array_equal(R.rowptr, At.rowptr) and np.array_equal(R.col, At.col)
    assert np.array_equal(R.val, At.val)


def test_general_csr_transpose_matches_reference():
    if O.ref_lib() is None:
        pytest.skip("reference SparseCPU.h driver not built")
    rng = np.random.default_rng(0)
    M = sp.random(50, 50, density=0.1, random_state=1, format="csr")
    A = O.Csr(50, M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data)
    R = O.ref_transpose(A)
    T = O.transpose(A)
    assert np.array_equal(R.rowptr, T.rowptr) and np.array_equal(R.col, T.col)
    assert np.array_equal(R.val, T.val)
    assert rng is not None


def test_jacobi_precondition_matches_reference():
    if O.ref_lib() is None:
        pytest.skip("reference SparseCPU.h driver not built")
    s = mesh.random_ldu(100, 300, seed=6)
    A = O.ldu2csr(s)
    inv = O.ref_jacobi_precondition(A)
    assert np.array_equal(inv, 1.0 / s.diag)


# ---------------------------------------------------------------------------
# SpMV and reductions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["hex", "random"])
def test_spmv_bit_exact_vs_scipy_and_close_to_ldu_amul(kind):
    s = mesh.laplacian(12, seed=3, symmetric=False) if kind == "hex" else \
        mesh.random_ldu(500, 2000, seed=7)
    A = O.ldu2csr(s)
    x = np.random.default_rng(1).standard_normal(s.n_cells)
    y = O.spmv(A, x)
    assert np.array_equal(y, A.scipy() @ x)   # sequential, unfused row sums
    y2 = mesh.ldu_matvec(s, x)
    np.testing.assert_allclose(y, y2, rtol=1e-13, atol=1e-13)
    r = O.residual(A, s.b, x)
    assert np.array_equal(r, s.b - y)


def test_spmv_f32_against_scipy_float32():
    s = mesh.laplacian(10, seed=1)
    A = O.ldu2csr(s)
    x = np.random.default_rng(2).standard_normal(s.n_cells).astype(np.float32)
    A32 = sp.csr_matrix((A.val.astype(np.float32), A.col, A.rowptr), shape=(A.n, A.n))
    assert np.array_equal(O.spmv_f32(A, x), A32 @ x)


def test_reduction_known_answers_pin_the_tree():
    # ulp(1e16) = 2, so 1e16 + 1 rounds back to 1e16 (ties to even) while 1e16 + 2,
    # + 4, + 6 are exact.  [1e16, 1 x 7] gives:
    #   leaves of 4 then a tree:  (((1e16+1)+1)+1) + (((1+1)+1)+1) = 1e16 + 4
    #   one sequential leaf of 8: every 1 is lost                  = 1e16
    #   a pure pairwise tree:     ((1e16+1)+(1+1)) + ((1+1)+(1+1)) = 1e16 + 6
    v = np.array([1e16] + [1.0] * 7)
    assert O.vsum(v) == 1e16 + 4
    # leaves 1e16, 8, -1e16 -> (1e16 + 8) + (-1e16) = 8; sequential summation gives 11
    w = np.array([1e16, 1, 1, 1, 2, 2, 2, 2, -1e16, 1, 1, 1], dtype=np.float64)
    assert O.vsum(w) == 8.0
    # dot products take the same tree (fma leaves): a.b with b = 1 equals the sum
    assert O.dot(v, np.ones(8)) == 1e16 + 4
    # tree over leaves: tree_sum([a, b, c]) = (a + b) + c
    assert O.tree_sum(np.array([1e16, 1.0, -1e16])) == 0.0
    assert O.tree_sum(np.array([1.0, 1e16, -1e16, 1.0])) == 0.0


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1000, 4097, 100003])
def test_dot_within_fsum_bound(n):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    exact = math.fsum((a * b).tolist())   # products differ from fma by < ulp each
    bound = 2 * (n + 4) * np.finfo(float).eps * float(np.sum(np.abs(a * b))) + 1e-300
    assert abs(O.dot(a, b) - exact) <= bound
    ints = rng.integers(-1000, 1000, n).astype(float)
    assert O.vsum(ints) == float(ints.sum())
    assert O.dot(ints, ints) == float((ints * ints).sum())
    assert O.dist(a, a) == 0.0
    assert abs(O.dist(a, b) - np.linalg.norm(a - b)) <= 1e-12 * np.linalg.norm(a - b)


def test_dot_f32_is_exact_product_fp64_accumulation():
    rng = np.random.default_rng(5)
    a = rng.standard_normal(777).astype(np.float32)
    b = rng.standard_normal(777).astype(np.float32)
    exact = math.fsum((a.astype(float) * b.astype(float)).tolist())
    assert abs(O.dot_f32(a, b) - exact) <= 1e-13 * float(np.sum(np.abs(a.astype(float) * b)))


# ---------------------------------------------------------------------------
# NormFactor (gis/gis.cu:106-123)
# ---------------------------------------------------------------------------
def test_norm_factor_closed_forms():
    s = mesh.laplacian(8, seed=2)
    A = O.ldu2csr(s)
    # x = 0: x_bar = 0, A x = A x_bar = 0  ->  nf = ||b|| + 1e-20
    nf0 = O.norm_factor(A, s.b, np.zeros(s.n_cells))
    assert abs(nf0 - (np.linalg.norm(s.b) + 1e-20)) <= 1e-14 * nf0
    # constant x = c: A x = A x_bar  ->  nf = ||b - A c1|| + 1e-20
    c = np.full(s.n_cells, 0.25)
    nfc = O.norm_factor(A, s.b, c)
    assert abs(nfc - np.linalg.norm(s.b - A.scipy() @ c)) <= 1e-13 * nfc
    # general x against an independent evaluation (LDU face loop, numpy norms)
    x = np.random.default_rng(3).uniform(-1, 1, s.n_cells)
    xb = np.full(s.n_cells, x.mean())
    ref = np.linalg.norm(mesh.ldu_matvec(s, x) - mesh.ldu_matvec(s, xb)) + \
        np.linalg.norm(s.b - mesh.ldu_matvec(s, xb)) + 1e-20
    assert abs(O.norm_factor(A, s.b, x) - ref) <= 1e-12 * ref


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k", [1, 5, 40])
def test_cg_matches_scipy_cg(k):
    s = mesh.laplacian(10, seed=4, shift=1e-2)
    A = O.ldu2csr(s)
    x0 = np.random.default_rng(0).uniform(-1, 1, s.n_cells)
    res = O.solve("CG", A, s.b, x0, 0.0, k)
    xs, _ = spla.cg(A.scipy(), s.b, x0=x0.copy(), rtol=0.0, atol=0.0, maxiter=k)
    assert res.iterations == k
    np.testing.assert_allclose(res.x, xs, rtol=1e-10, atol=1e-12)
    # recursive residual norms equal the true residual norms up to rounding
    assert abs(res.hist[-1] - np.linalg.norm(s.b - A.scipy() @ res.x)) <= 1e-9 * res.hist[0]


def test_cg_finite_termination_and_diagonal_one_step():
    s = mesh.laplacian(2, seed=1)   # N = 8 SPD
    A = O.ldu2csr(s)
    res = O.solve("CG", A, s.b, s.x0, 0.0, 8)
    assert res.hist[-1] <= 1e-10 * res.hist[0]
    xd = np.linalg.solve(_dense(A), s.b)
    np.testing.assert_allclose(res.x, xd, rtol=1e-9)
    N = 50
    d = np.linspace(2.0, 2.0, N)
    D = mesh.LduSystem(N, np.zeros(0, np.int32), np.zeros(0, np.int32), d, np.zeros(0),
                       np.zeros(0), np.random.default_rng(1).uniform(-1, 1, N), np.zeros(N))
    AD = O.ldu2csr(D)
    for m in ["CG", "Jacobi", "MR", "MRJacobi", "JMR", "MRS", "BiCG"]:
        r = O.solve(m, AD, D.b, D.x0, 0.0, 1)
        assert r.hist[1] <= 1e-15 * r.hist[0], m


def test_bicg_equals_cg_bitwise_on_symmetric_matrices():
    s = mesh.laplacian(9, seed=5, shift=1e-2)
    A = O.ldu2csr(s)
    cg = O.solve("CG", A, s.b, s.x0, 1e-9, 500)
    bi = O.solve("BiCG", A, s.b, s.x0, 1e-9, 500)
    assert np.array_equal(cg.hist, bi.hist) and np.array_equal(cg.x, bi.x)


@pytest.mark.parametrize("k", [1, 7, 30])
def test_bicg_matches_scipy_bicg(k):
    s = mesh.laplacian(8, seed=6, shift=0.1, symmetric=False)
    A = O.ldu2csr(s)
    res = O.solve("BiCG", A, s.b, s.x0, 0.0, k)
    xs, _ = spla.bicg(A.scipy(), s.b, x0=np.zeros(s.n_cells), rtol=0.0, atol=0.0, maxiter=k)
    np.testing.assert_allclose(res.x, xs, rtol=1e-9, atol=1e-12)


def test_bicg_breakdown():
    N = 16
    D = mesh.LduSystem(N, np.zeros(0, np.int32), np.zeros(0, np.int32), np.full(N, 1e11),
                       np.zeros(0), np.zeros(0), np.ones(N), np.zeros(N))
    r = O.solve("BiCG", O.ldu2csr(D), D.b, D.x0, 1e-30, 5)
    assert r.status == O.BREAKDOWN and r.hist_len == 1 and np.all(r.x == 0)


def test_jacobi_equals_classic_sweep():
    s = mesh.laplacian(6, seed=7, shift=0.5, symmetric=False)
    A = O.ldu2csr(s)
    Ad = _dense(A)
    D = np.diag(Ad)
    x = np.zeros(s.n_cells)
    for _ in range(15):
        x = (s.b - (Ad - np.diag(D)) @ x) / D
    res = O.solve("Jacobi", A, s.b, s.x0, 0.0, 15)
    np.testing.assert_allclose(res.x, x, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("m", ["MR", "MRJacobi", "MRS", "Jacobi", "Richardson", "CG", "BiCG"])
def test_recursive_residual_matches_true_residual(m):
    s = mesh.laplacian(8, seed=8, shift=0.3, symmetric=m in ("CG", "Richardson"))
    A = O.ldu2csr(s)
    res = O.solve(m, A, s.b, s.x0, 0.0, 25)
    true = np.linalg.norm(s.b - A.scipy() @ res.x)
    assert abs(res.hist[-1] - true) <= 1e-10 * res.hist[0]


def test_jmr_tracks_scaled_residual_and_mrs_equals_mrjacobi():
    s = mesh.laplacian(8, seed=9, shift=0.3, symmetric=False)
    A = O.ldu2csr(s)
    res = O.solve("JMR", A, s.b, s.x0, 0.0, 25)
    S = np.sqrt(1.0 / s.diag)
    true = np.linalg.norm(S * (s.b - A.scipy() @ res.x))
    assert abs(res.hist[-1] - true) <= 1e-10 * res.hist[0]
    a = O.solve("MRS", A, s.b, s.x0, 0.0, 25)
    b = O.solve("MRJacobi", A, s.b, s.x0, 0.0, 25)
    assert np.array_equal(a.hist, b.hist) and np.array_equal(a.x, b.x)


@pytest.mark.parametrize("m", ["MR", "MRJacobi", "JMR"])
def test_minimal_residual_monotone(m):
    s = mesh.laplacian(8, seed=10, shift=0.05, symmetric=False)
    res = O.solve(m, O.ldu2csr(s), s.b, s.x0, 0.0, 60)
    assert np.all(np.diff(res.hist) <= 1e-12 * res.hist[0])


@pytest.mark.parametrize("m", ["MR", "MRJacobi", "MRS", "JMR"])
def test_minimal_residual_step_is_the_least_squares_minimiser(m):
    """Each step of the MR family moves x along its direction d by the omega that
    minimises the (scaled) residual norm over the line: omega* = (r, W) / (W, W)
    with W the residual's change per unit step (SolverSingleGPU.h:825-890,
    :1000-1073, :1079-1155, :1236-1303), recomputed here in numpy from the true
    residual.  A wrong numerator, denominator, sign or direction fails it."""
    s = mesh.laplacian(7, seed=12, shift=0.2, symmetric=False)
    A = O.ldu2csr(s)
    As = A.scipy()
    D = s.diag
    S = 1.0 / np.sqrt(D)
    x0 = s.x0.copy()
    for k in range(1, 7):
        x1 = O.solve(m, A, s.b, s.x0, 0.0, k).x
        r = s.b - As @ x0
        if m == "MR":
            rr, d = r, r
            W = As @ d
        elif m in ("MRJacobi", "MRS"):
            rr, d = r, r / D
            W = As @ d
        else:   # JMR: residual scaled by D^-1/2, direction D^-1/2 of it
            rr = S * r
            d = S * rr
            W = S * (As @ d)
        omega_star = (rr @ W) / (W @ W)
        step = x1 - x0
        omega = (step @ d) / (d @ d)
        assert abs(omega - omega_star) <= 1e-9 * abs(omega_star), (k, omega, omega_star)
        assert np.linalg.norm(step - omega * d) <= 1e-10 * np.linalg.norm(step), k
        # and it is a minimum: nearby steps along d leave a larger residual
        base = np.linalg.norm(rr - omega_star * W)
        for e in (1e-3, -1e-3):
            assert np.linalg.norm(rr - omega_star * (1 + e) * W) > base
        x0 = x1


def test_richardson_power_estimate():
    s = mesh.laplacian(6, seed=11, shift=0.1)
    A = O.ldu2csr(s)
    lmax = np.linalg.eigvalsh(_dense(A)).max()
    omega = O.optimize_power(A, 30)
    lmin = O.find_diag_abs_min(A)
    assert lmin == np.abs(s.diag).min()
    lest = 2.0 / omega - lmin
    assert 0.8 * lmax <= lest <= lmax * (1 + 1e-12)   # a Rayleigh quotient
    # A = c I: lambda_max = lambda_min = c, omega = 1/c, one-step convergence
    N, c = 30, 4.0
    D = mesh.LduSystem(N, np.zeros(0, np.int32), np.zeros(0, np.int32), np.full(N, c),
                       np.zeros(0), np.zeros(0), np.random.default_rng(2).uniform(-1, 1, N),
                       np.zeros(N))
    r = O.solve("Richardson", O.ldu2csr(D), D.b, D.x0, 0.0, 1)
    assert abs(r.omega - 1.0 / c) <= 1e-15
    assert r.hist[1] <= 1e-14 * r.hist[0]


def test_foam_wrapper_semantics():
    s = mesh.laplacian(8, seed=12, shift=1e-2)
    A = O.ldu2csr(s)
    psi = np.random.default_rng(3).uniform(-1, 1, s.n_cells)
    x, perf = O.foam_solve("CG", s, psi.copy(), s.b, 1e-6, 1000, A=A)
    nf = O.norm_factor(A, s.b, psi)
    assert abs(perf.initial_residual - np.linalg.norm(s.b - A.scipy() @ psi)) <= \
        1e-12 * perf.initial_residual
    assert perf.final_residual <= 1e-6 * nf
    res = O.solve("CG", A, s.b, psi, 1e-6 * nf, 1000)
    assert perf.n_iterations == res.iterations and res.hist[-2] > 1e-6 * nf


# ---------------------------------------------------------------------------
# OpenFOAM's relTol / minIter (lduMatrix::solver; not passed on by the
# reference wrappers): pinned against a brute-force stop predictor.  Every
# method is deterministic, so a run with more iterations extends the same
# residual history; the stop index of any (error, relTol, minIter, maxIter) is
# then read off one long history.
# ---------------------------------------------------------------------------
def _predict_stop(hist, method, error, rel_tol, min_iter, max_iter):
    thr = max(error, rel_tol * hist[0]) if rel_tol > 0 else error
    conv = (lambda v: v < thr) if method == "BiCG" else (lambda v: not v > thr)
    for k in range(0 if method != "BiCG" else 1, len(hist)):
        if (conv(hist[k]) or k >= max_iter) and k >= min_iter:
            return k
    raise AssertionError("history too short for the prediction")


CONTROLS = [(1e-30, 0.1, 0, 500), (1e-30, 1e-3, 4, 500), (1e30, 0.0, 6, 500),
            (1e-30, 0.0, 9, 5), (1e-4, 0.5, 0, 500), (1e-30, 0.0, 0, 7)]


@pytest.mark.parametrize("method", list(O.METHODS))
@pytest.mark.parametrize("ctl", CONTROLS)
def test_rel_tol_and_min_iter_stop_where_predicted(method, ctl):
    error, rel_tol, min_iter, max_iter = ctl
    sym = method in ("CG", "Richardson")
    s = mesh.laplacian(7, seed=21, shift=0.3, symmetric=sym)
    A = O.ldu2csr(s)
    long = O.solve(method, A, s.b, s.x0, 0.0, 400)
    assert long.status == 0 and long.hist_len == 401
    k = _predict_stop(long.hist, method, error, rel_tol, min_iter, max_iter)
    r = O.solve(method, A, s.b, s.x0, error, max_iter, rel_tol=rel_tol, min_steps=min_iter)
    assert r.iterations == k
    np.testing.assert_array_equal(r.hist, long.hist[:k + 1])


def test_zero_controls_are_the_reference_loops():
    s = mesh.laplacian(7, seed=22, shift=0.3, symmetric=False)
    A = O.ldu2csr(s)
    for m in O.METHODS:
        if m in ("CG", "Richardson"):
            continue
        a = O.solve(m, A, s.b, s.x0, 1e-6, 200)
        b = O.solve(m, A, s.b, s.x0, 1e-6, 200, rel_tol=0.0, min_steps=0)
        np.testing.assert_array_equal(a.x, b.x)
        np.testing.assert_array_equal(a.hist, b.hist)


def test_foam_controls_rel_tol_counts_from_the_initial_residual():
    s = mesh.laplacian(8, seed=12, shift=1e-2)
    A = O.ldu2csr(s)
    psi = np.random.default_rng(3).uniform(-1, 1, s.n_cells)
    x, perf = O.foam_solve("CG", s, psi.copy(), s.b, 1e-12, 1000, A=A, rel_tol=1e-2)
    assert perf.final_residual <= 1e-2 * perf.initial_residual
    res = O.solve("CG", A, s.b, psi, 0.0, perf.n_iterations)
    assert res.hist[-2] > 1e-2 * perf.initial_residual
    _, perf2 = O.foam_solve("CG", s, psi.copy(), s.b, 1e-12, 1000, A=A, min_iter=3)
    _, perf3 = O.foam_solve("CG", s, psi.copy(), s.b, 1e-12, 1000, A=A)
    assert perf2.n_iterations == perf3.n_iterations > 3


# ---------------------------------------------------------------------------
# mixed precision (no reference counterpart): pinned by properties
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [0, 1])
def test_mixed_reaches_fp64_target(mode):
    s = mesh.laplacian(16, seed=13, shift=1e-4)
    A = O.ldu2csr(s)
    x, hist, it, st = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 5000, mode)
    assert st == 0
    true = np.linalg.norm(s.b - A.scipy() @ x) / np.linalg.norm(s.b)
    assert true <= 1e-10
    assert abs(hist[-1] / np.linalg.norm(s.b) - true) <= 1e-3 * true
    assert np.all(np.diff(hist) < 0)


def test_mixed_without_updates_tracks_fp64_cg_early():
    s = mesh.laplacian(16, seed=14, shift=1e-2)
    A = O.ldu2csr(s)
    x32, _, it, st = O.mixed_cg(A, s.b, s.x0, 1e-30, 0.0, 20, 0)
    assert it == 20 and st == 11
    r64 = O.solve("CG", A, s.b, s.x0, 0.0, 20)
    t32 = np.linalg.norm(s.b - A.scipy() @ x32)
    t64 = np.linalg.norm(s.b - A.scipy() @ r64.x)
    assert abs(t32 - t64) <= 1e-4 * t64


def test_openfoam_pcg_baseline_is_cg():
    """The CPU baseline's OpenFOAM PCG (preconditioner none) runs the same Krylov
    recurrence as the oracle's CG: equal iterates up to summation order at a fixed
    iteration count, and the normFactor closed form for x0 = 0 (nf = sum|b|, so the
    first scaled residual is 1)."""
    s = mesh.laplacian(10, seed=4, shift=1e-2)
    A = O.ldu2csr(s)
    x, it, res, _ = O.pcg_openfoam(s, 0.0, 25)
    ref = O.solve("CG", A, s.b, s.x0, 0.0, 25)
    assert it == 25
    np.testing.assert_allclose(x, ref.x, rtol=1e-9, atol=1e-12)
    r = s.b - O.spmv(A, x)
    assert res == pytest.approx(np.abs(r).sum() / np.abs(s.b).sum(), rel=1e-9)
    x0, it0, res0, _ = O.pcg_openfoam(s, 2.0, 100)     # tolerance above the first residual
    assert it0 == 0 and res0 == pytest.approx(1.0, rel=1e-15)


@pytest.mark.parametrize("method,symmetric", [("CG", True), ("BiCG", False)])
def test_mixed_krylov_refinement_reaches_fp64_target(method, symmetric):
    """Restarted FP32/FP64 refinement: every outer step shrinks the FP64 residual by
    about the inner tolerance, and the final x meets the FP64 target recomputed here."""
    s = mesh.laplacian(12, seed=5, shift=1e-2, symmetric=symmetric)
    A = O.ldu2csr(s)
    x, hist, inner, st = O.mixed_krylov(method, A, s.b, rel_tol=1e-12, inner_rel_tol=1e-3)
    assert st == 0
    bn = np.linalg.norm(s.b)
    assert np.linalg.norm(s.b - O.spmv(A, x)) <= 1e-12 * bn
    assert hist[-1] <= 1e-12 * bn < hist[-2]
    assert all(hist[j + 1] < 0.05 * hist[j] for j in range(len(hist) - 1))
    assert inner >= len(hist) - 1


def test_mixed_krylov_single_outer_step_is_an_fp32_solve():
    """One correction from x0 = 0 with a tight inner tolerance: x equals the FP32 Krylov
    solution, i.e. close to A^-1 b at single-precision accuracy, and the outer cap is
    reported."""
    s = mesh.laplacian(6, seed=1, shift=0.5, symmetric=True)
    A = O.ldu2csr(s)
    x, hist, inner, st = O.mixed_krylov("CG", A, s.b, rel_tol=1e-14, inner_rel_tol=1e-6,
                                        max_outer=1)
    assert st == 11 and len(hist) == 2
    xe = np.linalg.solve(mesh.ldu_to_dense(s), s.b)
    assert np.linalg.norm(x - xe) <= 1e-5 * np.linalg.norm(xe)
    assert hist[1] < 1e-4 * hist[0]
