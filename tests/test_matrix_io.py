"""Matrix/vector files of cuFoam's CPU types (gis/SparseSystem/SparseCPU.h) and
CSR -> LDU conversion.

Pinned three ways: against the reference's own SparseCPU.h readers/writers
(compiled from source into oracle/_ref), against the dense matrices the
fixtures were written from (tests/golden/make_matrix_files.py), and by round
trips through the oracle's ldu2csr."""
import os

import numpy as np
import pytest

from cufoam_amd import matrix_io as mio
from cufoam_amd import mesh
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RB = ["grid6_rsa", "grid5_rua", "rect4x6_rra", "skew5_rza"]


def _ref():
    if O.ref_lib() is None:
        pytest.skip("the reference's SparseCPU.h is not built here (oracle/_ref)")


def _dense():
    return np.load(os.path.join(GOLD, "matrix_files_dense.npz"))


@pytest.mark.parametrize("name", RB)
def test_read_rb_matches_reference_reader(name):
    _ref()
    path = os.path.join(GOLD, name + ".rb")
    A = mio.read(path)
    nr, nc, rp, col, val = O.ref_csr_read(path)
    assert (A.num_rows, A.num_cols) == (nr, nc)
    np.testing.assert_array_equal(A.row_offsets, rp)
    np.testing.assert_array_equal(A.column_indices, col)
    np.testing.assert_array_equal(A.values, val)


@pytest.mark.parametrize("name", RB)
def test_read_rb_reproduces_the_written_matrix(name):
    A = mio.read_rb(os.path.join(GOLD, name + ".rb"))
    D = _dense()[name]
    if name.endswith("_rza"):
        D = D.T   # readRB's skew expansion yields the transpose of the stored matrix
    np.testing.assert_array_equal(A.to_dense(), D)
    # rows come out with ascending columns
    for i in range(A.num_rows):
        c = A.column_indices[A.row_offsets[i]:A.row_offsets[i + 1]]
        assert np.all(np.diff(c) > 0)


@pytest.mark.parametrize("name", ["grid6_psa", "grid6_rse"])
def test_read_rb_rejects_what_the_reference_rejects(name):
    path = os.path.join(GOLD, name + ".rb")
    with pytest.raises(RuntimeError):
        mio.read_rb(path)
    if O.ref_lib() is not None:
        assert O.ref_csr_read(path) is None


def test_csr_binary_round_trip_and_reference_compatibility(tmp_path):
    A = mio.read_rb(os.path.join(GOLD, "grid5_rua.rb"))
    p1 = str(tmp_path / "a.csr")
    mio.save_csr(p1, A)
    B = mio.load_csr(p1)
    for f in ("row_offsets", "column_indices", "values"):
        np.testing.assert_array_equal(getattr(A, f), getattr(B, f))
    assert (B.num_rows, B.num_cols) == (A.num_rows, A.num_cols)
    if O.ref_lib() is None:
        return
    nr, nc, rp, col, val = O.ref_csr_read(p1)        # our writer -> reference reader
    np.testing.assert_array_equal(rp, A.row_offsets)
    np.testing.assert_array_equal(col, A.column_indices)
    np.testing.assert_array_equal(val, A.values)
    p2 = str(tmp_path / "b.bin")                     # reference writer -> our reader
    O.ref_csr_save(p2, A.num_rows, A.num_cols, A.row_offsets, A.column_indices, A.values)
    C = mio.read(p2)
    np.testing.assert_array_equal(C.row_offsets, A.row_offsets)
    np.testing.assert_array_equal(C.values, A.values)
    assert open(p1, "rb").read() == open(p2, "rb").read()


def test_csr_binary_errors(tmp_path):
    with pytest.raises(RuntimeError):
        mio.load_csr(str(tmp_path / "missing.csr"))
    p = str(tmp_path / "short.csr")
    np.array([3, 3, 9], np.int32).tofile(p)
    with pytest.raises(RuntimeError):
        mio.load_csr(p)


def test_vectors_against_reference(tmp_path):
    v = np.random.default_rng(1).standard_normal(11)
    p = str(tmp_path / "v.vec")
    mio.save_vector(p, v)
    np.testing.assert_array_equal(mio.load_vector(p), v)
    m = mio.read_vector_mtx(os.path.join(GOLD, "vec7.mtx"))
    np.testing.assert_array_equal(m, np.load(os.path.join(GOLD, "vec7.npy")))
    if O.ref_lib() is None:
        return
    np.testing.assert_array_equal(O.ref_vec_read(p, False), v)
    p2 = str(tmp_path / "w.vec")
    O.ref_vec_save(p2, v)
    assert open(p, "rb").read() == open(p2, "rb").read()
    np.testing.assert_array_equal(O.ref_vec_read(os.path.join(GOLD, "vec7.mtx"), True), m)


@pytest.mark.parametrize("symmetric", [True, False])
def test_csr_to_ldu_inverts_ldu2csr(symmetric):
    s = mesh.laplacian(6, seed=2, shift=0.1, symmetric=symmetric)
    A = O.ldu2csr(s)
    L = mio.csr_to_ldu(mio.SparseCSR(A.n, A.n, A.rowptr, A.col, A.val))
    np.testing.assert_array_equal(L.lower_addr, s.lower_addr)
    np.testing.assert_array_equal(L.upper_addr, s.upper_addr)
    np.testing.assert_array_equal(L.diag, s.diag)
    np.testing.assert_array_equal(L.lower, s.lower)
    np.testing.assert_array_equal(L.upper, s.upper)


def test_csr_to_ldu_fills_structural_zeros():
    # stored: a(0,0) a(0,2) | a(1,0) | a(2,1) a(2,2); faces (0,1) (0,2) (1,2), each with
    # one side missing, and no a(1,1)
    A = mio.SparseCSR(3, 3, np.array([0, 2, 3, 5], np.int32), np.array([0, 2, 0, 1, 2], np.int32),
                      np.array([4.0, -1.0, -2.0, -3.0, 5.0]))
    L = mio.csr_to_ldu(A)
    np.testing.assert_array_equal(L.lower_addr, [0, 0, 1])
    np.testing.assert_array_equal(L.upper_addr, [1, 2, 2])
    np.testing.assert_array_equal(L.upper, [0.0, -1.0, 0.0])
    np.testing.assert_array_equal(L.lower, [-2.0, 0.0, -3.0])
    np.testing.assert_array_equal(L.diag, [4.0, 0.0, 5.0])
    np.testing.assert_array_equal(mesh.ldu_to_dense(L), A.to_dense())
    C = O.ldu2csr(L)
    assert C.nnz == 9   # four explicit zeros: a(0,1), a(2,0), a(1,2), a(1,1)
    with pytest.raises(ValueError):
        mio.csr_to_ldu(mio.SparseCSR(2, 3, np.array([0, 1, 2], np.int32),
                                     np.array([0, 1], np.int32), np.ones(2)))
    with pytest.raises(ValueError):
        mio.csr_to_ldu(mio.SparseCSR(2, 2, np.array([0, 2, 2], np.int32),
                                     np.array([1, 1], np.int32), np.ones(2)))


@pytest.mark.parametrize("name", ["grid6_rsa", "grid5_rua"])
def test_rb_matrix_through_ldu_is_the_same_csr(name):
    A = mio.read_rb(os.path.join(GOLD, name + ".rb"))
    L = mio.csr_to_ldu(A)
    C = O.ldu2csr(L)
    np.testing.assert_array_equal(C.rowptr, A.row_offsets)
    np.testing.assert_array_equal(C.col, A.column_indices)
    np.testing.assert_array_equal(C.val, A.values)


def test_renumbered_mesh_is_the_permuted_matrix():
    """mesh.renumber: A' = P A P^T with the faces back in OpenFOAM order."""
    from cufoam_amd import mesh
    from oracle import oracle as O
    s = mesh.laplacian(5, seed=3, shift=0.2, symmetric=False)
    perm = np.random.default_rng(4).permutation(s.n_cells)
    r = mesh.renumber(s, perm)
    assert np.all(r.lower_addr < r.upper_addr)
    key = r.lower_addr.astype(np.int64) * s.n_cells + r.upper_addr
    assert np.all(np.diff(key) > 0)
    D = mesh.ldu_to_dense(s)
    P = np.zeros((s.n_cells, s.n_cells))
    P[perm, np.arange(s.n_cells)] = 1.0
    np.testing.assert_array_equal(mesh.ldu_to_dense(r), P @ D @ P.T)
    np.testing.assert_array_equal(r.b, P @ s.b)
    A = O.ldu2csr(r)
    np.testing.assert_allclose(A.scipy() @ (P @ s.b), P @ (D @ s.b), rtol=1e-13)
