"""Host-side mirror of cuFoam's OpenFOAM plugin surface (lduMatrix::solver).

cuFoam registers eight ``lduMatrix::solver`` subclasses that OpenFOAM selects
by name from ``fvSolution`` (``solver cuFoamCG; tolerance ...; maxIter ...;``).
This module keeps the same names, constructor arguments and ``solve``
contract over the gis C ABI, so a caller (or a test) drives the library
exactly as OpenFOAM drives the wrappers:

    solver = lduMatrix.solver.New("cuFoamCG", "p", matrix, controls)
    perf = solver.solve(psi, source)        # psi updated in place

Semantics reproduced from the reference (cuFoamCG/cuFoamCG.C:38-89 and the
seven siblings):
  * ``tolerance`` is absolute on the L2 residual, scaled by NormFactor
    (gis/gis.cu:106-123); ``maxIter`` is the iteration cap (<= 0 unbounded);
  * ``relTol``, ``minIter`` and ``cmpt`` are accepted and ignored, as there,
    unless the controls hold ``convergence openfoam``: then relTol and minIter
    act as in OpenFOAM's lduMatrix::solver (gis_foam_solve_controls);
  * the returned residuals are the unnormalised L2 norms history.front() and
    history.back(), nIterations = history.size() - 1;
  * cuFoamCG and cuFoamRichardson are registered for symmetric matrices only
    (cuFoamCG.C:14-15, cuFoamRichardson.C:11-12); the others for both.

The coupling arguments (interfaces, coupleBouCoeffs) are accepted and, as in
the reference, not used: processor and cyclic patches are not coupled.  With
``interfaces coupled`` in the controls, cyclic interfaces whose both sides are
cells of this mesh (``cyclicInterface``) become matrix entries
A(faceCells, nbrFaceCells) = -coupleBouCoeffs (gis_set_interfaces), the sign
of OpenFOAM's interface update in Amul.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .gis import Gis, Performance


@dataclass
class cyclicInterface:
    """One half of a cyclic patch pair: its face cells and, face by face, the
    cells on the other half (cyclicLduInterface::faceCells / neighbour faceCells)."""
    faceCells: np.ndarray
    nbrFaceCells: np.ndarray


@dataclass
class lduMatrix:
    """Minimal lduMatrix: coefficients plus lduAddressing (owner/neighbour)."""
    diag: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    lowerAddr: np.ndarray
    upperAddr: np.ndarray

    @classmethod
    def from_system(cls, sys) -> "lduMatrix":
        return cls(sys.diag, sys.lower, sys.upper, sys.lower_addr, sys.upper_addr)

    def symmetric(self) -> bool:
        return bool(np.array_equal(self.lower, self.upper))

    def asymmetric(self) -> bool:
        return not self.symmetric()

    def size(self) -> int:
        return int(self.diag.shape[0])


# name -> (gis method, registered for symmetric, registered for asymmetric)
_TABLE = {
    "cuFoamCG": ("CG", True, False),
    "cuFoamBiCG": ("BiCG", True, True),
    "cuFoamJacobi": ("Jacobi", True, True),
    "cuFoamRichardson": ("Richardson", True, False),
    "cuFoamMR": ("MR", True, True),
    "cuFoamMRJacobi": ("MRJacobi", True, True),
    "cuFoamJMR": ("JMR", True, True),
    "cuFoamMRS": ("MRS", True, True),
    # no counterpart in the reference: CG with a preconditioner (controls key
    # "preconditioner": none / diagonal / GAMG), for symmetric matrices
    "cuFoamPCG": ("PCG", True, False),
    # no counterpart in the reference: BiCG with a preconditioner (none /
    # diagonal / DILU), for symmetric and asymmetric matrices
    "cuFoamPBiCG": ("PBiCG", True, True),
}


class FatalError(RuntimeError):
    pass


class cuFoamSolver:
    """Base of the eight wrappers (one per TypeName)."""

    typeName = ""

    def __init__(self, fieldName: str, matrix: lduMatrix, coupleBouCoeffs=None,
                 coupleIntCoeffs=None, interfaces=None, solverControls: dict | None = None,
                 gis: Gis | None = None):
        controls = dict(solverControls or {})
        self.fieldName = fieldName
        self.matrix = matrix
        self.tolerance = float(controls.get("tolerance", 1e-6))
        self.relTol = float(controls.get("relTol", 0.0))        # ignored, as in the reference
        self.maxIter = int(controls.get("maxIter", 1000))
        self.minIter = int(controls.get("minIter", 0))          # ignored, as in the reference
        self.convergence = str(controls.get("convergence", "reference"))
        if self.convergence not in ("reference", "openfoam"):
            raise FatalError(f"convergence must be reference or openfoam, not "
                             f"{self.convergence}")
        self.preconditioner = str(controls.get("preconditioner", "none"))
        if self.preconditioner not in ("none", "diagonal", "DIC", "DILU", "GAMG", "GAMG32"):
            raise FatalError(f"preconditioner must be none, diagonal, DIC, DILU, GAMG or GAMG32, "
                             f"not {self.preconditioner}")
        if self.typeName == "cuFoamPBiCG" and self.preconditioner not in ("none", "diagonal",
                                                                          "DILU"):
            raise FatalError(f"cuFoamPBiCG takes the none, diagonal or DILU preconditioner, not "
                             f"{self.preconditioner}")
        if self.typeName == "cuFoamPCG" and self.preconditioner == "DILU":
            raise FatalError("cuFoamPCG takes none, diagonal, DIC, GAMG or GAMG32; DILU is for "
                             "asymmetric matrices (cuFoamPBiCG)")
        self.couple = str(controls.get("interfaces", "ignored"))
        if self.couple not in ("ignored", "coupled"):
            raise FatalError(f"interfaces must be ignored or coupled, not {self.couple}")
        self._cpl = None
        if self.couple == "coupled" and interfaces:
            rows, cols, bou = [], [], []
            for i, itf in enumerate(interfaces):
                if itf is None:
                    continue
                if not isinstance(itf, cyclicInterface):
                    raise FatalError("only cyclic interfaces within this mesh can be coupled")
                rows.append(np.asarray(itf.faceCells, np.int32))
                cols.append(np.asarray(itf.nbrFaceCells, np.int32))
                bou.append(np.asarray(coupleBouCoeffs[i], np.float64))
            if rows:
                self._cpl = (np.concatenate(rows), np.concatenate(cols), np.concatenate(bou))
        self._gis = gis if gis is not None else Gis()
        self._pattern_key = None

    def _upload(self):
        # a reused address with another mesh behind it differs in its sizes
        key = (id(self.matrix.lowerAddr), self.matrix.size(), len(self.matrix.lowerAddr),
               len(self.matrix.upperAddr), None if self._cpl is None else id(self._cpl[0]))
        if key != self._pattern_key:
            self._gis.set_pattern(self.matrix.size(), self.matrix.lowerAddr,
                                  self.matrix.upperAddr)
            if self._cpl is not None:
                self._gis.set_interfaces(self._cpl[0], self._cpl[1])
            self._pattern_key = key
        if self._cpl is not None:
            self._gis.set_coupled_values(self.matrix.diag, self.matrix.lower, self.matrix.upper,
                                         -self._cpl[2])
        else:
            self._gis.set_values(self.matrix.diag, self.matrix.lower, self.matrix.upper)

    def solve(self, psi: np.ndarray, source: np.ndarray, cmpt: int = 0) -> Performance:
        if psi.dtype != np.float64 or not psi.flags.c_contiguous:
            raise FatalError("psi must be a contiguous float64 field")
        if psi.size != self.matrix.size() or np.size(source) != self.matrix.size():
            raise FatalError(f"psi and source must have {self.matrix.size()} cells")
        self._upload()
        src = np.ascontiguousarray(source, dtype=np.float64)
        if self.typeName in ("cuFoamPCG", "cuFoamPBiCG"):
            of = self.convergence == "openfoam"
            perf = self._gis.foam_solve(_TABLE[self.typeName][0], psi, src, self.tolerance,
                                        self.maxIter,
                                        rel_tol=self.relTol if of else 0.0,
                                        min_iter=self.minIter if of else 0,
                                        preconditioner=self.preconditioner)
        elif self.convergence == "openfoam":
            perf = self._gis.foam_solve(_TABLE[self.typeName][0], psi, src, self.tolerance,
                                        self.maxIter, rel_tol=self.relTol,
                                        min_iter=self.minIter)
        else:
            perf = self._gis.foam_solve(_TABLE[self.typeName][0], psi, src, self.tolerance,
                                        self.maxIter)
        perf.solver = self.typeName
        return perf


def _make(name: str):
    return type(name, (cuFoamSolver,), {"typeName": name})


cuFoamCG = _make("cuFoamCG")
cuFoamBiCG = _make("cuFoamBiCG")
cuFoamJacobi = _make("cuFoamJacobi")
cuFoamRichardson = _make("cuFoamRichardson")
cuFoamMR = _make("cuFoamMR")
cuFoamMRJacobi = _make("cuFoamMRJacobi")
cuFoamJMR = _make("cuFoamJMR")
cuFoamMRS = _make("cuFoamMRS")
cuFoamPCG = _make("cuFoamPCG")
cuFoamPBiCG = _make("cuFoamPBiCG")

_CLASSES = {c.typeName: c for c in (cuFoamCG, cuFoamBiCG, cuFoamJacobi, cuFoamRichardson,
                                      cuFoamMR, cuFoamMRJacobi, cuFoamJMR, cuFoamMRS,
                                      cuFoamPCG, cuFoamPBiCG)}


def New(name: str, fieldName: str, matrix: lduMatrix, solverControls: dict,
        gis: Gis | None = None, interfaceBouCoeffs=None, interfaceIntCoeffs=None,
        interfaces=None) -> cuFoamSolver:
    """lduMatrix::solver::New: runtime selection by name and matrix symmetry."""
    if name not in _TABLE:
        raise FatalError(f"Unknown {'symmetric' if matrix.symmetric() else 'asymmetric'} "
                         f"matrix solver {name}; valid: {sorted(_TABLE)}")
    _, sym_ok, asym_ok = _TABLE[name]
    if matrix.symmetric() and not sym_ok or matrix.asymmetric() and not asym_ok:
        kind = "symmetric" if matrix.symmetric() else "asymmetric"
        raise FatalError(f"Unknown {kind} matrix solver {name}")
    return _CLASSES[name](fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                          solverControls, gis=gis)
