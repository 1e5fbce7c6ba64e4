"""Coupled interfaces on one device (no counterpart in the reference, whose
wrappers ignore interfaces, cuFoamCG/cuFoamCG.C:30-31): the oracle's coupled
CSR and its transpose, pinned on the CPU.

  * a periodic mesh built with its wrap-around faces as couplings must give the
    same dense matrix as the same mesh with those faces as ordinary LDU faces,
    and as OpenFOAM's face loop plus interface update (mesh.ldu_matvec);
  * rows keep the reference's CSR entries first (ldu2csr of the mesh without
    couplings), then the couplings in entry order;
  * mirror_transpose equals the dense transpose, equals the reference-order
    transpose (SparseCPU.h:369-399) on uncoupled meshes, and rejects patterns
    that are not structurally symmetric;
  * the oracle's solvers on the coupled CSR converge to the dense solution.
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O


def _dense(A: O.Csr) -> np.ndarray:
    return A.scipy().toarray()


@pytest.mark.parametrize("n,symmetric", [(2, True), (3, False), (6, True), (6, False)])
def test_coupled_csr_is_the_periodic_matrix(n, symmetric):
    s = mesh.periodic_laplacian(n, seed=n, shift=0.1, symmetric=symmetric)
    A = O.ldu2csr(s)
    D = mesh.ldu_to_dense(s)
    np.testing.assert_array_equal(_dense(A), D)
    if n > 2:   # at n = 2 the wrap faces coincide with internal faces
        F = mesh.couplings_as_faces(s)
        np.testing.assert_array_equal(_dense(O.ldu2csr(F)), D)
    x = np.random.default_rng(1).standard_normal(s.n_cells)
    np.testing.assert_allclose(O.spmv(A, x), D @ x, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(mesh.ldu_matvec(s, x), D @ x, rtol=1e-13, atol=1e-12)


def test_coupled_rows_append_entries_after_the_reference_row():
    s = mesh.periodic_laplacian(5, seed=2, shift=0.1, symmetric=False)
    A = O.ldu2csr(s)
    plain = mesh.laplacian(5, seed=2, symmetric=False)
    P = O.ldu2csr(mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, s.lower,
                                 s.upper, s.b, s.x0))
    assert np.array_equal(plain.lower_addr, s.lower_addr)
    for r in range(s.n_cells):
        a, b = A.rowptr[r], A.rowptr[r + 1]
        pa, pb = P.rowptr[r], P.rowptr[r + 1]
        np.testing.assert_array_equal(A.col[a:a + pb - pa], P.col[pa:pb])
        np.testing.assert_array_equal(A.val[a:a + pb - pa], P.val[pa:pb])
        js = np.nonzero(s.cpl_row == r)[0]
        np.testing.assert_array_equal(A.col[a + pb - pa:b], s.cpl_col[js])
        np.testing.assert_array_equal(A.val[a + pb - pa:b], s.cpl_val[js])


@pytest.mark.parametrize("n", [2, 4, 7])
def test_mirror_transpose(n):
    s = mesh.periodic_laplacian(n, seed=3, shift=0.1, symmetric=False)
    A = O.ldu2csr(s)
    T = O.mirror_transpose(A)
    np.testing.assert_array_equal(T.rowptr, A.rowptr)
    np.testing.assert_array_equal(T.col, A.col)
    np.testing.assert_array_equal(_dense(T), _dense(A).T)
    u = mesh.laplacian(n, seed=3, symmetric=False)
    U = O.ldu2csr(u)
    R = O.transpose(U)
    V = O.mirror_transpose(U)
    for f in ("rowptr", "col", "val"):
        np.testing.assert_array_equal(getattr(V, f), getattr(R, f))
    with pytest.raises(ValueError):
        O.mirror_transpose(O.append_entries(U, [0], [n * n * n - 1], [1.0]))


@pytest.mark.parametrize("method", ["CG", "BiCG", "MRJacobi", "Jacobi"])
def test_oracle_solves_the_coupled_system(method):
    sym = method == "CG"
    s = mesh.periodic_laplacian(6, seed=4, shift=0.5, symmetric=sym)
    A = O.ldu2csr(s)
    At = O.mirror_transpose(A) if method == "BiCG" else None
    r = O.solve(method, A, s.b, s.x0, 1e-11 * np.linalg.norm(s.b), 5000, At=At)
    xd = np.linalg.solve(mesh.ldu_to_dense(s), s.b)
    np.testing.assert_allclose(r.x, xd, rtol=1e-8, atol=1e-9)


def test_coupled_matrix_differs_from_the_ignored_interfaces():
    s = mesh.periodic_laplacian(4, seed=5, shift=0.1)
    A = O.ldu2csr(s)
    u = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, s.lower, s.upper, s.b,
                       s.x0)
    assert A.nnz == O.ldu2csr(u).nnz + s.n_couples
    x = np.ones(s.n_cells)
    assert not np.array_equal(O.spmv(A, x), O.spmv(O.ldu2csr(u), x))


def test_foam_interface_control_is_validated_before_any_gpu_call():
    from cufoam_amd import foam
    s = mesh.periodic_laplacian(3, seed=0, shift=0.1)
    m = foam.lduMatrix.from_system(s)
    with pytest.raises(foam.FatalError):
        foam.New("cuFoamCG", "p", m, {"interfaces": "implicit"})
    itf = [object()]
    with pytest.raises(foam.FatalError):
        foam.New("cuFoamCG", "p", m, {"interfaces": "coupled"}, interfaceBouCoeffs=[[1.0]],
                 interfaces=itf)
