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

    @property
    def n_faces(self) -> int:
        return int(self.lower_addr.shape[0])

    @property
    def nnz(self) -> int:
        return self.n_cells + 2 * self.n_faces

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


def batch_edges(count: int = 64, lo: int = 16, hi: int = 64, seed: int = 0) -> np.ndarray:
    """Mesh edges of an uneven batch, n_s ~ U{lo..hi}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count)


def ldu_matvec(sys: LduSystem, x: np.ndarray) -> np.ndarray:
    """y = A x by OpenFOAM's face loop (``lduMatrix::Amul`` without interfaces)."""
    y = sys.diag * x
    np.add.at(y, sys.upper_addr, sys.lower * x[sys.lower_addr])
    np.add.at(y, sys.lower_addr, sys.upper * x[sys.upper_addr])
    return y


def ldu_to_dense(sys: LduSystem) -> np.ndarray:
    A = np.diag(sys.diag.astype(np.float64))
    A[sys.lower_addr, sys.upper_addr] = sys.upper
    A[sys.upper_addr, sys.lower_addr] = sys.lower
    return A
