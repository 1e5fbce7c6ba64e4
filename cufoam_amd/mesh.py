"""Synthetic OpenFOAM-style LDU systems on structured hex meshes.

An ``LduSystem`` holds exactly what an OpenFOAM ``lduMatrix`` hands to a
``lduMatrix::solver``: ``diag[N]``, ``lower[M]``, ``upper[M]`` and the face
addressing ``lowerAddr[M]`` (owner) / ``upperAddr[M]`` (neighbour), plus a
right-hand side ``b`` and an initial guess ``x0``.

Conventions (OpenFOAM ``lduAddressing``, the layout ``gis/ldu2csr.h:19-39``
consumes):

* A(owner, neighbour) = upper[f],  A(neighbour, owner) = lower[f];
* faces are in upper-triangular order: sorted by owner, then by neighbour.

Mesh: n^3 cells in lexicographic order c = i + n*j + n^2*k.  For c ascending
the internal faces (c, c+1) if i < n-1, (c, c+n) if j < n-1 and
(c, c+n^2) if k < n-1 are emitted, giving M = 3 n^2 (n-1) faces.

Coefficients (all draws from ``numpy.random.default_rng(seed)``, PCG64):

* symmetric: u ~ U[0,1) (M draws, face order); upper = lower = -(1 + 0.5 u);
* nonsymmetric (convection-diffusion like): upper = -(1 + 0.5 u) and
  lower = upper * (1 + 0.5 v), v ~ U[0,1) (M draws after u);
* diag_i = (sum of |off-diagonals| of row i) + shift;
* b ~ U[-1, 1] (N draws after the coefficients, one block per right-hand side).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class LduSystem:
    n_cells: int
    lower_addr: np.ndarray  # int32[M], owner of each face
    upper_addr: np.ndarray  # int32[M], neighbour of each face
    diag: np.ndarray        # float64[N]
    lower: np.ndarray       # float64[M]
    upper: np.ndarray       # float64[M]
    b: np.ndarray           # float64[N] or float64[nrhs, N]
    x0: np.ndarray          # float64[N] or float64[nrhs, N]
    meta: dict = field(default_factory=dict)
    # coupled interfaces (cyclic patches): extra entries A(cpl_row[j], cpl_col[j]) =
    # cpl_val[j], i.e. -coupleBouCoeffs of the patch face; None when uncoupled
    cpl_row: np.ndarray | None = None
    cpl_col: np.ndarray | None = None
    cpl_val: np.ndarray | None = None

    @property
    def n_faces(self) -> int:
        return int(self.lower_addr.shape[0])

    @property
    def n_couples(self) -> int:
        return 0 if self.cpl_row is None else int(self.cpl_row.shape[0])

    @property
    def nnz(self) -> int:
        return self.n_cells + 2 * self.n_faces + self.n_couples

    @property
    def symmetric(self) -> bool:
        return bool(np.array_equal(self.lower, self.upper))


def hex_faces(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Owner/neighbour addressing of the n^3 structured mesh, upper-triangular order."""
    if n < 1:
        raise ValueError("mesh edge must be >= 1")
    N = n * n * n
    c = np.arange(N, dtype=np.int64)
    i = c % n
    j = (c // n) % n
    k = c // (n * n)
    nbr = np.stack([c + 1, c + n, c + n * n], axis=1)
    ok = np.stack([i < n - 1, j < n - 1, k < n - 1], axis=1)
    owner = np.broadcast_to(c[:, None], nbr.shape)[ok]
    neigh = nbr[ok]
    return owner.astype(np.int32), neigh.astype(np.int32)


def row_abs_sum(n_cells: int, lower_addr, upper_addr, lower, upper) -> np.ndarray:
    """sum_j |a_ij| over the off-diagonals of every row (np.bincount sums in face order)."""
    s_up = np.bincount(lower_addr, weights=np.abs(upper), minlength=n_cells)
    s_lo = np.bincount(upper_addr, weights=np.abs(lower), minlength=n_cells)
    return s_up + s_lo


def laplacian(n: int, seed: int = 0, shift: float = 1e-3, symmetric: bool = True,
              nrhs: int = 1) -> LduSystem:
    """7-point FV Laplacian (symmetric) or convection-diffusion (nonsymmetric) on n^3 cells."""
    owner, neigh = hex_faces(n)
    N = n * n * n
    M = owner.shape[0]
    rng = np.random.default_rng(seed)
    u = rng.random(M)
    upper = -(1.0 + 0.5 * u)
    if symmetric:
        lower = upper.copy()
    else:
        v = rng.random(M)
        lower = upper * (1.0 + 0.5 * v)
    diag = row_abs_sum(N, owner, neigh, lower, upper) + shift
    if nrhs == 1:
        b = rng.uniform(-1.0, 1.0, N)
        x0 = np.zeros(N)
    else:
        b = rng.uniform(-1.0, 1.0, (nrhs, N))
        x0 = np.zeros((nrhs, N))
    return LduSystem(N, owner, neigh, diag, lower, upper, b, x0,
                     meta={"n": n, "seed": seed, "shift": shift, "symmetric": symmetric,
                           "nrhs": nrhs})


def periodic_laplacian(n: int, seed: int = 0, shift: float = 1e-3,
                       symmetric: bool = True) -> LduSystem:
    """laplacian(n) on a mesh that is periodic in x: the faces between i = n-1 and
    i = 0 are a cyclic patch pair, held as interface entries rather than faces.

    The internal faces and their coefficients are those of laplacian(n, seed,
    symmetric=...).  Then, per cyclic face (j, k ascending), w ~ U[0,1) and (for
    the nonsymmetric case) v ~ U[0,1) give the coefficient -(1 + 0.5 w) of the
    half at i = 0 (row (0,j,k), column (n-1,j,k)) and -(1 + 0.5 w)(1 + 0.5 v) of the
    half at i = n-1.  Entries list the i = 0 half first, then the i = n-1 half,
    as OpenFOAM orders a cyclic pair's interfaces.  diag includes |coupling|
    terms, so the matrix stays diagonally dominant; b is drawn last.
    """
    if n < 2:
        raise ValueError("a periodic mesh needs n >= 2")
    s = laplacian(n, seed, shift=shift, symmetric=symmetric)
    rng = np.random.default_rng([seed, 7919])
    jk = np.arange(n * n)
    c0 = (n * jk).astype(np.int32)            # cells (0, j, k)
    c1 = (n * jk + n - 1).astype(np.int32)    # cells (n-1, j, k)
    w = -(1.0 + 0.5 * rng.random(n * n))
    w1 = w.copy() if symmetric else w * (1.0 + 0.5 * rng.random(n * n))
    row = np.concatenate([c0, c1])
    col = np.concatenate([c1, c0])
    val = np.concatenate([w, w1])
    diag = row_abs_sum(s.n_cells, s.lower_addr, s.upper_addr, s.lower, s.upper)
    diag = diag + np.bincount(row, weights=np.abs(val), minlength=s.n_cells) + shift
    s.diag = diag
    s.b = rng.uniform(-1.0, 1.0, s.n_cells)
    s.cpl_row, s.cpl_col, s.cpl_val = row, col, val
    s.meta = dict(s.meta, periodic="x")
    return s


def couplings_as_faces(sys: LduSystem) -> LduSystem:
    """The same matrix with each mirrored pair of coupling entries made an ordinary
    face (owner = lower cell index), faces re-sorted to upper-triangular order.
    Needs at most one coupling entry per (row, col)."""
    r, c, v = sys.cpl_row, sys.cpl_col, sys.cpl_val
    keys = {(int(a), int(b)): float(x) for a, b, x in zip(r, c, v)}
    if len(keys) != len(r):
        raise ValueError("duplicate coupling entries")
    own, nei, up, lo = [], [], [], []
    for (a, b), x in keys.items():
        if a < b:
            own.append(a)
            nei.append(b)
            up.append(x)
            lo.append(keys[(b, a)])
    la = np.concatenate([sys.lower_addr, np.array(own, np.int32)])
    ua = np.concatenate([sys.upper_addr, np.array(nei, np.int32)])
    lower = np.concatenate([sys.lower, np.array(lo)])
    upper = np.concatenate([sys.upper, np.array(up)])
    order = np.lexsort((ua, la))
    return LduSystem(sys.n_cells, la[order].astype(np.int32), ua[order].astype(np.int32),
                     sys.diag.copy(), lower[order], upper[order], sys.b, sys.x0,
                     meta=dict(sys.meta, couplings="as faces"))


def config_system(name: str, seed: int = 0) -> LduSystem:
    """The five synthetic workloads, by name.

    * ``poisson32``   32^3 SPD, shift 1e-3           (CPU plumbing case)
    * ``poisson128``  128^3 SPD, shift 1e-3          (FP64 drop-in)
    * ``convdiff128`` 128^3 nonsymmetric, shift 0.1  (FP64 drop-in, BiCG / MRJacobi)
    * ``poisson256``  256^3 SPD, shift 1e-4          (mixed precision)
    * ``convdiff128x3`` 128^3 nonsymmetric, 3 right-hand sides
    * ``poisson64``   64^3 SPD, shift 1e-3           (adjoint)
    """
    table = {
        "poisson32": dict(n=32, shift=1e-3, symmetric=True),
        "poisson64": dict(n=64, shift=1e-3, symmetric=True),
        "poisson128": dict(n=128, shift=1e-3, symmetric=True),
        "convdiff128": dict(n=128, shift=0.1, symmetric=False),
        "poisson256": dict(n=256, shift=1e-4, symmetric=True),
        "convdiff128x3": dict(n=128, shift=0.1, symmetric=False, nrhs=3),
    }
    if name not in table:
        raise KeyError(f"unknown workload {name!r}; known: {sorted(table)}")
    return laplacian(seed=seed, **table[name])


def random_ldu(n_cells: int, n_faces: int, seed: int = 0, symmetric: bool = False,
               shift: float = 0.5) -> LduSystem:
    """Unstructured LDU system: n_faces distinct random cell pairs (owner < neighbour)
    in upper-triangular order (sorted by owner, then neighbour), random coefficients,
    diagonally dominant diagonal.  Rows have irregular lengths, so neighbour groups
    (losort) interleave owners in a way structured meshes never show."""
    rng = np.random.default_rng(seed)
    max_pairs = n_cells * (n_cells - 1) // 2
    n_faces = min(n_faces, max_pairs)
    pairs = set()
    while len(pairs) < n_faces:
        a = rng.integers(0, n_cells, size=2 * (n_faces - len(pairs)) + 8)
        c = rng.integers(0, n_cells, size=a.shape[0])
        for u, v in zip(a.tolist(), c.tolist()):
            if u != v and len(pairs) < n_faces:
                pairs.add((min(u, v), max(u, v)))
    arr = np.array(sorted(pairs), dtype=np.int64).reshape(-1, 2)
    owner = arr[:, 0].astype(np.int32)
    neigh = arr[:, 1].astype(np.int32)
    M = owner.shape[0]
    upper = -(0.5 + rng.random(M))
    lower = upper.copy() if symmetric else -(0.5 + rng.random(M))
    diag = row_abs_sum(n_cells, owner, neigh, lower, upper) + shift
    b = rng.uniform(-1.0, 1.0, n_cells)
    return LduSystem(n_cells, owner, neigh, diag, lower, upper, b, np.zeros(n_cells),
                     meta={"random": True, "seed": seed, "symmetric": symmetric})


def renumber(sys: LduSystem, perm: np.ndarray) -> LduSystem:
    """The same system with cell c renamed perm[c] (an unstructured numbering of a
    structured mesh): faces re-oriented so owner < neighbour (upper and lower
    swap with the orientation), then sorted by owner and neighbour as OpenFOAM
    orders them; diag, b and x0 move with their cells."""
    perm = np.asarray(perm, np.int64)
    a = perm[sys.lower_addr]
    c = perm[sys.upper_addr]
    flip = a > c
    own = np.where(flip, c, a)
    nbr = np.where(flip, a, c)
    up = np.where(flip, sys.lower, sys.upper)   # A(own, nbr)
    lo = np.where(flip, sys.upper, sys.lower)   # A(nbr, own)
    order = np.lexsort((nbr, own))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.shape[0])
    return LduSystem(sys.n_cells, own[order].astype(np.int32), nbr[order].astype(np.int32),
                     sys.diag[inv], lo[order], up[order], sys.b[..., inv], sys.x0[..., inv],
                     meta=dict(sys.meta, renumbered=True))


def batch_edges(count: int = 64, lo: int = 16, hi: int = 64, seed: int = 0) -> np.ndarray:
    """Mesh edges of an uneven batch, n_s ~ U{lo..hi}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count)


def ldu_matvec(sys: LduSystem, x: np.ndarray) -> np.ndarray:
    """y = A x by OpenFOAM's face loop (``lduMatrix::Amul``), then the interface
    entries (the interface update subtracts coupleBouCoeffs * psi(neighbour))."""
    y = sys.diag * x
    np.add.at(y, sys.upper_addr, sys.lower * x[sys.lower_addr])
    np.add.at(y, sys.lower_addr, sys.upper * x[sys.upper_addr])
    if sys.n_couples:
        np.add.at(y, sys.cpl_row, sys.cpl_val * x[sys.cpl_col])
    return y


def ldu_to_dense(sys: LduSystem) -> np.ndarray:
    A = np.diag(sys.diag.astype(np.float64))
    A[sys.lower_addr, sys.upper_addr] = sys.upper
    A[sys.upper_addr, sys.lower_addr] = sys.lower
    if sys.n_couples:
        np.add.at(A, (sys.cpl_row, sys.cpl_col), sys.cpl_val)
    return A
