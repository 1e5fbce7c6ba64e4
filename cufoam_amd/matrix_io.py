"""Matrix and vector files of cuFoam's CPU types, and CSR -> LDU conversion.

File formats (gis/SparseSystem/SparseCPU.h), reproduced so that files move
between this package and the reference unchanged:

* CSR binary (``SparseCSR::save`` / ``load``, SparseCPU.h:105-146): int32
  ``num_rows, num_cols, num_entries``; float64 ``values[num_entries]``; int32
  ``row_offsets[0 .. num_rows)`` (the closing ``num_entries`` is implied); int32
  ``column_indices[num_entries]``.  Native byte order (little-endian here).
* Rutherford-Boeing text (``SparseCSR::readRB``, SparseCPU.h:164-297), real or
  integer, assembled (``??a``) matrices: unsymmetric / rectangular files are
  transposed from their column-compressed storage; symmetric and
  skew-symmetric ones are expanded from the stored triangle.
* ``read`` dispatches on the suffix: ``.rb`` -> Rutherford-Boeing, anything
  else -> CSR binary (SparseCPU.h:305-311).
* Vector binary (``Vector::save`` / ``load``, SparseCPU.h:653-660, :683-687):
  int32 ``size``, float64 ``values[size]``; Matrix Market array text
  (``Vector::readMTX``, SparseCPU.h:662-681).

``csr_to_ldu`` turns a square CSR matrix into the LDU arrays a gis handle takes
(faces = structurally nonzero pairs i < j in owner order, explicit zeros where
only one of a(i,j), a(j,i) is stored), so external matrices run through the
same device path as OpenFOAM's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import LduSystem


@dataclass
class SparseCSR:
    """CPU::SparseCSR<double, int>: row_offsets[num_rows + 1], column_indices and
    values[num_entries]."""
    num_rows: int
    num_cols: int
    row_offsets: np.ndarray
    column_indices: np.ndarray
    values: np.ndarray

    @property
    def num_entries(self) -> int:
        return int(self.values.shape[0])

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.num_rows, self.num_cols))
        rows = np.repeat(np.arange(self.num_rows), np.diff(self.row_offsets))
        np.add.at(d, (rows, self.column_indices), self.values)
        return d


def _csr(nr, nc, rp, ci, v) -> SparseCSR:
    return SparseCSR(int(nr), int(nc), np.ascontiguousarray(rp, np.int32),
                     np.ascontiguousarray(ci, np.int32), np.ascontiguousarray(v, np.float64))


# ---- CSR binary --------------------------------------------------------------------
def save_csr(path: str, A: SparseCSR) -> None:
    """SparseCSR::save (SparseCPU.h:134-146)."""
    with open(path, "wb") as f:
        np.array([A.num_rows, A.num_cols, A.num_entries], np.int32).tofile(f)
        np.ascontiguousarray(A.values, np.float64).tofile(f)
        np.ascontiguousarray(A.row_offsets[:A.num_rows], np.int32).tofile(f)
        np.ascontiguousarray(A.column_indices, np.int32).tofile(f)


def load_csr(path: str) -> SparseCSR:
    """SparseCSR::load (SparseCPU.h:105-127); raises RuntimeError for an unreadable
    or truncated file."""
    try:
        with open(path, "rb") as f:
            head = np.fromfile(f, np.int32, 3)
            if head.shape[0] != 3:
                raise RuntimeError(f"SparseCSR::load: {path} is too short")
            nr, nc, nnz = (int(v) for v in head)
            if nr < 0 or nc < 0 or nnz < 0:
                raise RuntimeError(f"SparseCSR::load: bad sizes in {path}")
            val = np.fromfile(f, np.float64, nnz)
            rp = np.fromfile(f, np.int32, nr)
            ci = np.fromfile(f, np.int32, nnz)
    except OSError as e:
        raise RuntimeError(f"SparseCSR::load can't read file {path}") from e
    if val.shape[0] != nnz or rp.shape[0] != nr or ci.shape[0] != nnz:
        raise RuntimeError(f"SparseCSR::load: {path} is truncated")
    return _csr(nr, nc, np.append(rp, np.int32(nnz)), ci, val)


# ---- Rutherford-Boeing -------------------------------------------------------------
def transpose(A: SparseCSR) -> SparseCSR:
    """SparseCSR::transpose (SparseCPU.h:369-399): counting sort by column, so the rows
    within each column of the result stay in ascending order."""
    rows = np.repeat(np.arange(A.num_rows, dtype=np.int64), np.diff(A.row_offsets))
    order = np.argsort(A.column_indices, kind="stable")
    counts = np.bincount(A.column_indices, minlength=A.num_cols)
    rp = np.zeros(A.num_cols + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    return _csr(A.num_cols, A.num_rows, rp, rows[order], A.values[order])


def _expand_triangle(A: SparseCSR, sign: float) -> SparseCSR:
    """Mirror the stored triangle into a full matrix in readRB's fill order
    (SparseCPU.h:227-296): for stored entry j of row i (in file order), row i gets
    (i, c, v) and, off the diagonal, row c gets (c, i, sign * v); entries of a row
    are placed in increasing j."""
    n = A.num_rows
    i = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.row_offsets))
    c = A.column_indices.astype(np.int64)
    v = A.values
    j = np.arange(A.num_entries, dtype=np.int64)
    off = i != c
    rows = np.concatenate([i, c[off]])
    cols = np.concatenate([c, i[off]])
    vals = np.concatenate([v, sign * v[off]])
    keys = np.concatenate([j, j[off]])
    order = np.lexsort((keys, rows))
    counts = np.bincount(rows, minlength=n)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    return _csr(n, n, rp, cols[order], vals[order])


def read_rb(path: str) -> SparseCSR:
    """SparseCSR::readRB (SparseCPU.h:164-297).

    Reads the type/size line (line 3) and the column-compressed arrays as
    whitespace-separated numbers, ignoring the Fortran formats of line 4.  Types:
    ``r``/``i`` values only (RuntimeError otherwise), assembled ``a`` only.
    ``u``/``r`` matrices are transposed into CSR; ``s`` matrices are expanded from
    their stored triangle; ``z`` (skew-symmetric) matrices are expanded as readRB
    does, with the stored value in the transposed position and its negative
    mirrored, which yields the transpose of the stored matrix.  The reference
    leaves the matrix unchanged for the symmetry letter ``h``; this reader raises
    RuntimeError for it.  Unlike the reference, the type letters are accepted in
    either case and Fortran ``D`` exponents are read.
    """
    try:
        with open(path, "r") as f:
            lines = f.read().splitlines()
    except OSError as e:
        raise RuntimeError(f"SparseCSR::readRB can't read file {path}") from e
    if len(lines) < 4:
        raise RuntimeError(f"SparseCSR::readRB: {path} has no header")
    head = lines[2].split()
    if len(head) < 4:
        raise RuntimeError(f"SparseCSR::readRB: bad type line in {path}")
    mtype = head[0].lower()
    if len(mtype) < 3 or mtype[0] not in "ri":
        raise RuntimeError("SparseCSR::readRB supports real matrices only")
    if mtype[2] != "a":
        raise RuntimeError("SparseCSR::readRB does not support the elemental format")
    rows, cols, entries = int(head[1]), int(head[2]), int(head[3])
    tok = " ".join(lines[4:]).replace("D", "E").replace("d", "e").split()
    need = cols + 1 + 2 * entries
    if len(tok) < need:
        raise RuntimeError(f"SparseCSR::readRB: {path} is truncated")
    ptr = np.array(tok[:cols + 1], dtype=np.int64) - 1
    idx = np.array(tok[cols + 1:cols + 1 + entries], dtype=np.int64) - 1
    val = np.array(tok[cols + 1 + entries:need], dtype=np.float64)
    At = _csr(cols, rows, ptr, idx, val)   # the CSC of A read as the CSR of A^T
    sym = mtype[1]
    if sym in "ur":
        return transpose(At)
    if sym == "s":
        return _expand_triangle(At, 1.0)
    if sym == "z":
        return _expand_triangle(At, -1.0)
    raise RuntimeError(f"SparseCSR::readRB: unsupported symmetry type {mtype!r}")


def read(path: str) -> SparseCSR:
    """SparseCSR::read (SparseCPU.h:305-311): ``.rb`` -> readRB, else load."""
    return read_rb(path) if path.endswith(".rb") else load_csr(path)


# ---- vectors -----------------------------------------------------------------------
def save_vector(path: str, v) -> None:
    """Vector::save (SparseCPU.h:683-687)."""
    v = np.ascontiguousarray(v, np.float64)
    with open(path, "wb") as f:
        np.array([v.shape[0]], np.int32).tofile(f)
        v.tofile(f)


def load_vector(path: str) -> np.ndarray:
    """Vector::load (SparseCPU.h:653-660)."""
    try:
        with open(path, "rb") as f:
            n = np.fromfile(f, np.int32, 1)
            if n.shape[0] != 1 or n[0] < 0:
                raise RuntimeError(f"Vector::load: bad size in {path}")
            v = np.fromfile(f, np.float64, int(n[0]))
    except OSError as e:
        raise RuntimeError(f"Vector::load can't read file {path}") from e
    if v.shape[0] != int(n[0]):
        raise RuntimeError(f"Vector::load: {path} is truncated")
    return v


def read_vector_mtx(path: str) -> np.ndarray:
    """Vector::readMTX (SparseCPU.h:662-681): line 1 skipped, the first number of
    line 2 is the size, the values follow one per token."""
    try:
        with open(path, "r") as f:
            lines = f.read().splitlines()
    except OSError as e:
        raise RuntimeError(f"Vector::readMTX can't read file {path}") from e
    if len(lines) < 2 or not lines[1].split():
        raise RuntimeError(f"Vector::readMTX: {path} has no size line")
    n = int(lines[1].split()[0])
    tok = " ".join(lines[2:]).split()
    if len(tok) < n:
        raise RuntimeError(f"Vector::readMTX: {path} is truncated")
    return np.array(tok[:n], dtype=np.float64)


# ---- CSR -> LDU ----------------------------------------------------------------------
def csr_to_ldu(A: SparseCSR, b=None, x0=None) -> LduSystem:
    """LDU arrays of a square CSR matrix.

    Faces are the structurally nonzero pairs (i, j), i < j, in OpenFOAM's
    upper-triangular order (by owner i, then neighbour j): upper[f] = a(i, j),
    lower[f] = a(j, i), each 0.0 where the CSR does not store it; diag[i] = a(i, i)
    or 0.0.  Duplicate (i, j) entries are rejected.  For a CSR with ascending
    columns in every row, a stored diagonal and a structurally symmetric pattern,
    the device pattern of the result is the input CSR itself, so the products
    (and every solve) use the same values in the same order.
    """
    if A.num_rows != A.num_cols:
        raise ValueError("csr_to_ldu needs a square matrix")
    n = A.num_rows
    i = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.row_offsets))
    j = A.column_indices.astype(np.int64)
    if np.any(j < 0) or np.any(j >= n):
        raise ValueError("column index out of range")
    key = i * n + j
    if np.unique(key).shape[0] != key.shape[0]:
        raise ValueError("csr_to_ldu: duplicate entries")
    diag = np.zeros(n)
    d = i == j
    diag[i[d]] = A.values[d]
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    off = ~d
    pair = np.unique(lo[off] * n + hi[off])          # sorted by owner, then neighbour
    owner = (pair // n).astype(np.int32)
    neigh = (pair % n).astype(np.int32)
    upper = np.zeros(pair.shape[0])
    lower = np.zeros(pair.shape[0])
    f = np.searchsorted(pair, lo[off] * n + hi[off])
    up = (i < j)[off]
    upper[f[up]] = A.values[off][up]
    lower[f[~up]] = A.values[off][~up]
    b = np.zeros(n) if b is None else np.ascontiguousarray(b, np.float64)
    x0 = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
    return LduSystem(n, owner, neigh, diag, lower, upper, b, x0,
                     meta={"from_csr": True, "nnz": A.num_entries})
