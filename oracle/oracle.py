"""ctypes front end of the CPU oracle (oracle/gis_oracle*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.

Every function returns fresh numpy arrays; inputs are not modified.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgis_oracle.so")
BASELINE_PATH = os.path.join(HERE, "_build", "libcpu_baseline.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_sparsecpu.so")

METHODS = {"CG": 0, "BiCG": 1, "Jacobi": 2, "MR": 3, "MRJacobi": 4, "JMR": 5, "MRS": 6,
           "Richardson": 7}
_PBICG_DIAG = 13   # or_solve_ctl's id for BiCG preconditioned by the diagonal
BREAKDOWN = 10
NAN = 12

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False) -> None:
    """Compile the oracle with its Makefile (gcc only; no GPU toolchain needed)."""
    if force or not os.path.exists(LIB_PATH) or not os.path.exists(BASELINE_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref() -> bool:
    """Compile the reference's SparseCPU.h driver when the cuFoam sources are present."""
    if not os.path.isdir("/root/reference/gis/SparseSystem"):
        return os.path.exists(REF_PATH)
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.or_dot.restype = C.c_double
        L.or_dot.argtypes = [C.c_int64, _f64p, _f64p]
        L.or_dot_f32.restype = C.c_double
        L.or_dot_f32.argtypes = [C.c_int64, _f32p, _f32p]
        L.or_sum.restype = C.c_double
        L.or_sum.argtypes = [C.c_int64, _f64p]
        L.or_dist.restype = C.c_double
        L.or_dist.argtypes = [C.c_int64, _f64p, _f64p]
        L.or_tree_sum.restype = C.c_double
        L.or_tree_sum.argtypes = [C.c_int64, _f64p]
        L.or_ldu2csr_pattern.restype = C.c_int
        L.or_ldu2csr_pattern.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.or_ldu2csr_values.restype = None
        L.or_ldu2csr_values.argtypes = [C.c_int, C.c_int, C.c_int64, _i32p, _f64p, _f64p, _f64p,
                                        _f64p]
        L.or_transpose.restype = None
        L.or_transpose.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]
        L.or_spmv.restype = None
        L.or_spmv.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_f32.restype = None
        L.or_spmv_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_residual.restype = None
        L.or_residual.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.or_norm_factor.restype = C.c_double
        L.or_norm_factor.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_find_diag_abs_min.restype = C.c_double
        L.or_find_diag_abs_min.argtypes = [C.c_int, _i32p, _i32p, _f64p]
        L.or_optimize_power.restype = C.c_double
        L.or_optimize_power.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int]
        L.or_solve_ctl.restype = C.c_int
        L.or_solve_ctl.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, _f64p, _f64p, C.c_double, C.c_int,
                                   C.c_double, C.c_int, C.c_double, C.c_double, C.c_int64, _f64p,
                                   C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.or_solve.restype = C.c_int
        L.or_solve.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_void_p, C.c_void_p,
                               C.c_void_p, _f64p, _f64p, C.c_double, C.c_int, C.c_double,
                               C.c_int, C.c_double, _f64p, C.c_int, C.POINTER(C.c_int),
                               C.POINTER(C.c_double)]
        L.or_mixed_krylov.restype = C.c_int
        L.or_mixed_krylov.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, _f64p, _f64p, C.c_double,
                                      C.c_double, C.c_int, C.c_int, _f64p, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_cg_par.restype = C.c_int
        L.or_cg_par.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                _f64p, C.c_int, C.POINTER(C.c_int)]
        L.or_mixed_cg_par.restype = C.c_int
        L.or_mixed_cg_par.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                      C.c_double, C.c_int, C.c_int, _f64p, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_residual_norm_par.restype = C.c_double
        L.or_residual_norm_par.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.or_par_threads.restype = C.c_int
        L.or_par_threads.argtypes = []
        if hasattr(L, "or_mixed_cg"):
            L.or_mixed_cg.restype = C.c_int
            L.or_mixed_cg.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                      C.c_double, C.c_int, C.c_int, _f64p, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# formats
# ---------------------------------------------------------------------------
@dataclass
class Csr:
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))


def ldu2csr_pattern(n_cells, lower_addr, upper_addr):
    """(rowptr, col, perm) of gis/ldu2csr.h:19-39; perm codes lower f / M+c diag / M+N+f upper."""
    la = _c(lower_addr, np.int32)
    ua = _c(upper_addr, np.int32)
    M = la.shape[0]
    nnz = n_cells + 2 * M
    rowptr = np.zeros(n_cells + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    perm = np.zeros(max(nnz, 1), np.int32)
    st = lib().or_ldu2csr_pattern(n_cells, M, la, ua, rowptr, col, perm)
    if st != 0:
        raise ValueError(f"or_ldu2csr_pattern failed with status {st}")
    return rowptr, col[:nnz], perm[:nnz]


def ldu2csr(sys) -> Csr:
    """CPU CSR of an LduSystem, as getSparseCSRMat builds it (gis/ldu2csr.h:7-42).

    Interface entries (sys.cpl_*; no counterpart in the reference, which drops
    them) follow each row's LDU entries, in entry order (append_entries)."""
    rowptr, col, perm = ldu2csr_pattern(sys.n_cells, sys.lower_addr, sys.upper_addr)
    nnz = perm.shape[0]
    val = np.zeros(max(nnz, 1))
    lib().or_ldu2csr_values(sys.n_cells, sys.n_faces, nnz, _c(perm if nnz else np.zeros(1), np.int32),
                            _c(sys.diag, np.float64), _c(sys.lower if sys.n_faces else np.zeros(1), np.float64),
                            _c(sys.upper if sys.n_faces else np.zeros(1), np.float64), val)
    A = Csr(sys.n_cells, rowptr, col, val[:nnz])
    if getattr(sys, "cpl_row", None) is not None and len(sys.cpl_row):
        A = append_entries(A, sys.cpl_row, sys.cpl_col, sys.cpl_val)
    return A


def append_entries(A: Csr, row, col, val) -> Csr:
    """A with extra entries (row[j], col[j], val[j]) placed after the existing
    entries of their row, in j order."""
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    all_rows = np.concatenate([rows, np.asarray(row, np.int64)])
    order = np.argsort(all_rows, kind="stable")
    counts = np.bincount(all_rows, minlength=A.n)
    rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    c = np.concatenate([A.col, np.asarray(col, np.int32)])[order].astype(np.int32)
    v = np.concatenate([A.val, np.asarray(val, np.float64)])[order]
    return Csr(A.n, rowptr, c, v)


def mirror_transpose(A: Csr) -> Csr:
    """A^T on A's own pattern, for a structurally symmetric A: the k-th entry of
    column c in row r takes the value of the k-th entry of column r in row c.
    (For an LDU matrix whose rows ascend in column this is transpose(A).)"""
    pos: dict[tuple[int, int], list[int]] = {}
    for r in range(A.n):
        for p in range(A.rowptr[r], A.rowptr[r + 1]):
            pos.setdefault((r, int(A.col[p])), []).append(p)
    val = np.empty_like(A.val)
    for (r, c), ps in pos.items():
        qs = pos.get((c, r))
        if qs is None or len(qs) != len(ps):
            raise ValueError("mirror_transpose needs a structurally symmetric pattern")
        val[ps] = A.val[qs]
    return Csr(A.n, A.rowptr.copy(), A.col.copy(), val)


def transpose(A: Csr) -> Csr:
    """A^T by the counting sort of SparseCPU.h:369-399."""
    t_rowptr = np.zeros(A.n + 1, np.int32)
    t_col = np.zeros(max(A.nnz, 1), np.int32)
    t_val = np.zeros(max(A.nnz, 1))
    lib().or_transpose(A.n, A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                       _c(A.val, np.float64), t_rowptr, t_col, t_val)
    return Csr(A.n, t_rowptr, t_col[:A.nnz], t_val[:A.nnz])


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def dot(a, b) -> float:
    a = _c(a, np.float64)
    return lib().or_dot(a.shape[0], a, _c(b, np.float64))


def dot_f32(a, b) -> float:
    a = _c(a, np.float32)
    return lib().or_dot_f32(a.shape[0], a, _c(b, np.float32))


def vsum(a) -> float:
    a = _c(a, np.float64)
    return lib().or_sum(a.shape[0], a)


def dist(a, b) -> float:
    a = _c(a, np.float64)
    return lib().or_dist(a.shape[0], a, _c(b, np.float64))


def tree_sum(leaves) -> float:
    leaves = _c(leaves, np.float64)
    return lib().or_tree_sum(leaves.shape[0], leaves)


def spmv(A: Csr, x) -> np.ndarray:
    y = np.zeros(A.n)
    lib().or_spmv(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32), _c(A.val, np.float64),
                  _c(x, np.float64), y)
    return y


def spmv_f32(A: Csr, x) -> np.ndarray:
    y = np.zeros(A.n, np.float32)
    lib().or_spmv_f32(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                      _c(A.val, np.float32), _c(x, np.float32), y)
    return y


def residual(A: Csr, b, x) -> np.ndarray:
    r = np.zeros(A.n)
    lib().or_residual(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32), _c(A.val, np.float64),
                      _c(b, np.float64), _c(x, np.float64), r)
    return r


def norm_factor(A: Csr, b, x) -> float:
    return lib().or_norm_factor(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                _c(A.val, np.float64), _c(b, np.float64), _c(x, np.float64))


def find_diag_abs_min(A: Csr) -> float:
    return lib().or_find_diag_abs_min(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                      _c(A.val, np.float64))


def optimize_power(A: Csr, steps: int = 30) -> float:
    return lib().or_optimize_power(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                   _c(A.val, np.float64), steps)


@dataclass
class SolveResult:
    x: np.ndarray
    hist: np.ndarray
    hist_len: int
    status: int
    omega: float = 0.0

    @property
    def iterations(self) -> int:
        return self.hist_len - 1


def solve(method: str, A: Csr, b, x0, error: float, max_steps: int, At: Csr | None = None,
          omega: float = 0.0, power_steps: int = 30, int_bound: float = 1e-10,
          hist_cap: int | None = None, rel_tol: float = 0.0,
          min_steps: int = 0) -> SolveResult:
    """One gis solver (SingleGPU::<method> semantics) on the CPU; x0 is copied.

    rel_tol / min_steps add OpenFOAM's relTol and minIter (lduMatrix::solver);
    0 and 0 are the reference's loops."""
    m = _PBICG_DIAG if method == "PBiCG-diagonal" else METHODS[method]
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    if hist_cap is None:
        hist_cap = max(max_steps if max_steps > 0 else 100000, min_steps) + 1
    hist = np.zeros(max(hist_cap, 1))
    hl = C.c_int(0)
    om = C.c_double(0.0)
    keep = []
    if m in (METHODS["BiCG"], _PBICG_DIAG):
        if At is None:
            At = transpose(A)
        keep = [_c(At.rowptr, np.int32), _c(At.col, np.int32), _c(At.val, np.float64)]
        tptrs = [k.ctypes.data for k in keep]
    else:
        tptrs = [None, None, None]
    st = lib().or_solve_ctl(m, A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                            _c(A.val, np.float64), tptrs[0], tptrs[1], tptrs[2],
                            _c(b, np.float64), x, error, max_steps, omega, power_steps,
                            int_bound, rel_tol, min_steps, hist, hist_cap, C.byref(hl),
                            C.byref(om))
    del keep
    if st not in (0, BREAKDOWN):
        raise RuntimeError(f"or_solve({method}) failed with status {st}")
    n = min(hl.value, hist_cap)
    if st == 0 and n == hl.value and n > 0 and np.isnan(hist[n - 1]):
        st = NAN   # same outcome code as the C ABI (GIS_NAN)
    return SolveResult(x, hist[:n].copy(), hl.value, st, om.value)


def pbicg(precond: str, A: Csr, b, x0, error: float, max_steps: int, At: Csr | None = None,
          int_bound: float = 1e-10, hist_cap: int | None = None, rel_tol: float = 0.0,
          min_steps: int = 0) -> SolveResult:
    """BiCG preconditioned by none (BiCG itself) or the diagonal (no counterpart
    in the reference; gis_oracle.c pbicg_diag)."""
    if precond not in ("none", "diagonal"):
        raise ValueError(f"PBiCG takes none or diagonal, not {precond}")
    return solve("BiCG" if precond == "none" else "PBiCG-diagonal", A, b, x0, error, max_steps,
                 At=At, int_bound=int_bound, hist_cap=hist_cap, rel_tol=rel_tol,
                 min_steps=min_steps)


@dataclass
class Performance:
    """The three numbers cuFoam's wrappers put into solverPerformance."""
    solver: str
    initial_residual: float
    final_residual: float
    n_iterations: int
    status: int


def foam_solve(method: str, sys, psi, source, tolerance: float, max_iter: int,
               A: Csr | None = None, rel_tol: float = 0.0, min_iter: int = 0,
               preconditioner: str = "none") -> tuple[np.ndarray, Performance]:
    """cuFoam<method>::solve on the CPU (cuFoamCG/cuFoamCG.C:45-83 and siblings).

    error = tolerance * NormFactor(A, b, x); the solver runs with max_steps = maxIter;
    initialResidual = hist.front(), finalResidual = hist.back(), nIterations = size-1.
    rel_tol / min_iter: OpenFOAM's relTol and minIter, which the reference
    wrappers do not pass on (0 and 0 reproduce them).  method "PCG" / "PBiCG":
    CG / BiCG with `preconditioner` (no counterpart in the reference).
    """
    if A is None:
        A = ldu2csr(sys)
    nf = norm_factor(A, source, psi)
    if method == "PCG":
        res = pcg(preconditioner, sys, source, psi, tolerance * nf, max_iter, rel_tol=rel_tol,
                  min_steps=min_iter)
        perf = Performance("cuFoamPCG", float(res.hist[0]), float(res.hist[-1]),
                           res.hist_len - 1, res.status)
        return res.x, perf
    At = None
    if method in ("BiCG", "PBiCG"):
        coupled = getattr(sys, "cpl_row", None) is not None and len(sys.cpl_row)
        At = mirror_transpose(A) if coupled else transpose(A)
    if method == "PBiCG" and preconditioner == "DILU":
        res = pbicg_ldu("DILU", sys, source, psi, tolerance * nf, max_iter, rel_tol=rel_tol,
                        min_steps=min_iter)
    elif method == "PBiCG":
        res = pbicg(preconditioner, A, source, psi, tolerance * nf, max_iter, At=At,
                    rel_tol=rel_tol, min_steps=min_iter)
    else:
        res = solve(method, A, source, psi, tolerance * nf, max_iter, At=At, rel_tol=rel_tol,
                    min_steps=min_iter)
    perf = Performance("cuFoam" + method, float(res.hist[0]), float(res.hist[-1]),
                       res.hist_len - 1, res.status)
    return res.x, perf


# ---------------------------------------------------------------------------
# preconditioned CG (none / diagonal / GAMG), gis_oracle_amg.c
# ---------------------------------------------------------------------------
PRECONDITIONERS = {"none": 0, "diagonal": 1, "GAMG": 2, "GAMG32": 3, "DIC": 4, "DILU": 5}
GAMG_OMEGA, GAMG_NU, GAMG_CSCALE = 0.8, 1, 1.8


def _amg_sigs(L):
    L.or_pcg.restype = C.c_int
    L.or_pcg.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                         _f64p, C.c_double, C.c_int, C.c_double, C.c_int64, C.c_double, C.c_int,
                         C.c_double, _f64p, C.c_int, C.POINTER(C.c_int)]
    L.or_amg_build.restype = C.c_void_p
    L.or_amg_build.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                               C.c_int, C.c_double]
    L.or_amg_free.argtypes = [C.c_void_p]
    L.or_amg_levels.argtypes = [C.c_void_p]
    L.or_amg_level_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.or_amg_level_get.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 7
    L.or_amg_dense_inverse.argtypes = [C.c_void_p, _f64p]
    L.or_amg_vcycle.argtypes = [C.c_void_p, _f64p, _f64p]
    L.or_amg_vcycle32.restype = C.c_int
    L.or_amg_vcycle32.argtypes = [C.c_void_p, _f64p, _f64p]
    L.or_pair_pass.restype = C.c_int
    L.or_pair_pass.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p]


def pair_pass(n: int, la, ua, w):
    """One pairwise agglomeration pass -> (aggregate count, cell -> aggregate)."""
    L = lib()
    if not hasattr(L, "_amg_sigs"):
        _amg_sigs(L)
        L._amg_sigs = True
    agg = np.zeros(n, np.int32)
    la = np.asarray(la, np.int32)
    nc = L.or_pair_pass(n, la.shape[0], _z(la, np.int32), _z(np.asarray(ua, np.int32), np.int32),
                        _z(np.asarray(w, np.float64), np.float64), agg)
    return nc, agg


def _z(a, dt):
    return _c(a if len(a) else np.zeros(1, dt), dt)


def dic_rd(sys) -> np.ndarray:
    """DIC's reciprocal factored diagonal (OpenFOAM's calcReciprocalD, face order)."""
    L = lib()
    L.or_dic_rd.restype = None
    L.or_dic_rd.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    rd = np.zeros(sys.n_cells)
    L.or_dic_rd(sys.n_cells, sys.n_faces, _z(sys.lower_addr, np.int32),
                _z(sys.upper_addr, np.int32), _c(sys.diag, np.float64),
                _z(sys.upper, np.float64), rd)
    return rd


def dic_apply(sys, rd, r) -> np.ndarray:
    """z = M^-1 r of DIC: z = rD r, forward face sweep, backward face sweep."""
    L = lib()
    L.or_dic_apply.restype = None
    L.or_dic_apply.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
    z = np.zeros(sys.n_cells)
    L.or_dic_apply(sys.n_cells, sys.n_faces, _z(sys.lower_addr, np.int32),
                   _z(sys.upper_addr, np.int32), _c(rd, np.float64), _z(sys.upper, np.float64),
                   _c(r, np.float64), z)
    return z


def dilu_rd(sys) -> np.ndarray:
    """DILU's reciprocal factored diagonal (rD[u] -= upper lower / rD[l], face order)."""
    L = lib()
    L.or_dilu_rd.restype = None
    L.or_dilu_rd.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
    rd = np.zeros(sys.n_cells)
    L.or_dilu_rd(sys.n_cells, sys.n_faces, _z(sys.lower_addr, np.int32),
                 _z(sys.upper_addr, np.int32), _c(sys.diag, np.float64),
                 _z(sys.lower, np.float64), _z(sys.upper, np.float64), rd)
    return rd


def dilu_apply(sys, rd, r, transpose: bool = False) -> np.ndarray:
    """z = M^-1 r (transpose: M^-T r) of DILU: z = rD r, forward and backward face sweeps."""
    L = lib()
    L.or_dilu_apply.restype = None
    L.or_dilu_apply.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int,
                                _f64p, _f64p]
    z = np.zeros(sys.n_cells)
    L.or_dilu_apply(sys.n_cells, sys.n_faces, _z(sys.lower_addr, np.int32),
                    _z(sys.upper_addr, np.int32), _c(rd, np.float64),
                    _z(sys.lower, np.float64), _z(sys.upper, np.float64), int(transpose),
                    _c(r, np.float64), z)
    return z


def pbicg_ldu(precond: str, sys, b, x0, error: float, max_steps: int,
              int_bound: float = 1e-10, rel_tol: float = 0.0, min_steps: int = 0,
              hist_cap: int | None = None) -> SolveResult:
    """BiCG preconditioned by the diagonal or DILU on an LduSystem
    (gis_oracle_amg.c or_pbicg); x0 is copied."""
    if precond not in ("diagonal", "DILU"):
        raise ValueError(f"pbicg_ldu takes diagonal or DILU, not {precond}")
    L = lib()
    L.or_pbicg.restype = C.c_int
    L.or_pbicg.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                           _f64p, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int64,
                           _f64p, C.c_int, C.POINTER(C.c_int)]
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    if hist_cap is None:
        hist_cap = max(max_steps if max_steps > 0 else 100000, min_steps) + 1
    hist = np.zeros(max(hist_cap, 1))
    hl = C.c_int(0)
    st = L.or_pbicg(PRECONDITIONERS[precond], sys.n_cells, sys.n_faces,
                    _z(sys.lower_addr, np.int32), _z(sys.upper_addr, np.int32),
                    _c(sys.diag, np.float64), _z(sys.lower, np.float64),
                    _z(sys.upper, np.float64), _c(b, np.float64), x, error, max_steps,
                    int_bound, rel_tol, min_steps, hist, hist_cap, C.byref(hl))
    if st not in (0, BREAKDOWN):
        raise RuntimeError(f"or_pbicg({precond}) failed with status {st}")
    n = min(hl.value, hist_cap)
    return SolveResult(x, hist[:n].copy(), hl.value, st)


def pcg(precond: str, sys, b, x0, error: float, max_steps: int, rel_tol: float = 0.0,
        min_steps: int = 0, omega: float = GAMG_OMEGA, nu: int = GAMG_NU,
        cscale: float = GAMG_CSCALE, hist_cap: int | None = None) -> SolveResult:
    """Preconditioned CG on an LduSystem (gis_oracle_amg.c); x0 is copied."""
    L = lib()
    if not hasattr(L, "_amg_sigs"):
        _amg_sigs(L)
        L._amg_sigs = True
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    if hist_cap is None:
        hist_cap = max(max_steps if max_steps > 0 else 100000, min_steps) + 1
    hist = np.zeros(max(hist_cap, 1))
    hl = C.c_int(0)
    st = L.or_pcg(PRECONDITIONERS[precond], sys.n_cells, sys.n_faces,
                  _z(sys.lower_addr, np.int32), _z(sys.upper_addr, np.int32),
                  _c(sys.diag, np.float64), _z(sys.lower, np.float64), _z(sys.upper, np.float64),
                  _c(b, np.float64), x, error, max_steps, rel_tol, min_steps, omega, nu, cscale,
                  hist, hist_cap, C.byref(hl))
    if st != 0:
        raise RuntimeError(f"or_pcg({precond}) failed with status {st}")
    n = min(hl.value, hist_cap)
    return SolveResult(x, hist[:n].copy(), hl.value, 0)


class Amg:
    """The GAMG hierarchy of an LduSystem (inspection and V-cycle)."""

    def __init__(self, sys, omega: float = GAMG_OMEGA, nu: int = GAMG_NU,
                 cscale: float = GAMG_CSCALE):
        L = lib()
        if not hasattr(L, "_amg_sigs"):
            _amg_sigs(L)
            L._amg_sigs = True
        self._L = L
        self.n = sys.n_cells
        self._h = L.or_amg_build(sys.n_cells, sys.n_faces, _z(sys.lower_addr, np.int32),
                                 _z(sys.upper_addr, np.int32), _c(sys.diag, np.float64),
                                 _z(sys.lower, np.float64), _z(sys.upper, np.float64), omega,
                                 nu, cscale)
        if not self._h:
            raise RuntimeError("or_amg_build failed")

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.or_amg_free(self._h)
            self._h = None

    @property
    def n_levels(self) -> int:
        return self._L.or_amg_levels(self._h)

    def level(self, l: int) -> dict:
        n, m, nc = C.c_int(0), C.c_int(0), C.c_int(0)
        self._L.or_amg_level_size(self._h, l, C.byref(n), C.byref(m), C.byref(nc))
        n, m, nc = n.value, m.value, nc.value
        out = {"n": n, "m": m, "nc": nc,
               "la": np.zeros(max(m, 1), np.int32), "ua": np.zeros(max(m, 1), np.int32),
               "diag": np.zeros(n), "lower": np.zeros(max(m, 1)), "upper": np.zeros(max(m, 1)),
               "agg": np.zeros(n, np.int32), "fcode": np.zeros(max(m, 1), np.int32)}
        self._L.or_amg_level_get(self._h, l, *[out[k].ctypes.data for k in
                                               ("la", "ua", "diag", "lower", "upper", "agg",
                                                "fcode")])
        for k in ("la", "ua", "lower", "upper", "fcode"):
            out[k] = out[k][:m]
        if nc == 0:
            out["agg"] = None
            out["fcode"] = None
        return out

    def dense_inverse(self):
        n = self.level(self.n_levels - 1)["n"]
        out = np.zeros(n * n)
        if self._L.or_amg_dense_inverse(self._h, out) != 0:
            return None
        return out.reshape(n, n)

    def vcycle(self, r):
        z = np.zeros(self.n)
        self._L.or_amg_vcycle(self._h, _c(r, np.float64), z)
        return z

    def vcycle32(self, r):
        """The FP32 cycle: (float) r in, (double) z out."""
        z = np.zeros(self.n)
        if self._L.or_amg_vcycle32(self._h, _c(r, np.float64), z) != 0:
            raise RuntimeError("or_amg_vcycle32 failed")
        return z


def mixed_cg(A: Csr, b, x0, rel_tol: float = 1e-10, inner_rel_tol: float = 1e-4,
             max_iter: int = 100000, mode: int = 0, hist_cap: int = 256):
    """Mixed-precision CG (FP32 inner Krylov, FP64 residual/correction/norms).

    mode 0: reliable updates (the Krylov direction survives each FP64 update);
    mode 1: restarted iterative refinement (inner CG restarted from zero).
    Returns (x, true-residual history, inner iterations, status).
    """
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    hist = np.zeros(hist_cap)
    hl = C.c_int(0)
    iters = C.c_int(0)
    st = lib().or_mixed_cg(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                           _c(A.val, np.float64), _c(b, np.float64), x, rel_tol, inner_rel_tol,
                           max_iter, mode, hist, hist_cap, C.byref(hl), C.byref(iters))
    return x, hist[:min(hl.value, hist_cap)].copy(), iters.value, st


# ---------------------------------------------------------------------------
# OpenMP restatements for the headline sizes (gis_oracle_par.c): the same bits
# as solve("CG", ...) and mixed_cg(...)
# ---------------------------------------------------------------------------
def par_threads() -> int:
    """OpenMP threads the parallel restatements use (OMP_NUM_THREADS)."""
    return lib().or_par_threads()


def cg_par(A: Csr, b, x0, error: float, max_steps: int, hist_cap: int | None = None):
    """CG (SolverSingleGPU.h:366-445) on all host threads -> SolveResult; bit for
    bit the result of solve("CG", A, b, x0, error, max_steps)."""
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    if hist_cap is None:
        hist_cap = (max_steps if max_steps > 0 else 100000) + 1
    hist = np.zeros(max(hist_cap, 1))
    hl = C.c_int(0)
    st = lib().or_cg_par(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                         _c(A.val, np.float64), _c(b, np.float64), x, error, max_steps, hist,
                         hist_cap, C.byref(hl))
    if st != 0:
        raise RuntimeError(f"or_cg_par failed with status {st}")
    return SolveResult(x, hist[:min(hl.value, hist_cap)].copy(), hl.value, 0)


def mixed_cg_par(A: Csr, b, x0, rel_tol: float = 1e-10, inner_rel_tol: float = 1e-4,
                 max_iter: int = 100000, mode: int = 0, hist_cap: int = 256):
    """mixed_cg on all host threads (same bits) -> (x, hist, inner iterations, status)."""
    x = _c(np.array(x0, dtype=np.float64, copy=True), np.float64)
    hist = np.zeros(hist_cap)
    hl = C.c_int(0)
    iters = C.c_int(0)
    st = lib().or_mixed_cg_par(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                               _c(A.val, np.float64), _c(b, np.float64), x, rel_tol,
                               inner_rel_tol, max_iter, mode, hist, hist_cap, C.byref(hl),
                               C.byref(iters))
    return x, hist[:min(hl.value, hist_cap)].copy(), iters.value, st


def residual_norm_par(A: Csr, b, x) -> float:
    """||b - A x||_2 in FP64 with the oracle's row sums and reduction tree."""
    r = np.empty(A.n)
    return lib().or_residual_norm_par(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                      _c(A.val, np.float64), _c(b, np.float64),
                                      _c(x, np.float64), r)


# ---------------------------------------------------------------------------
# the reference's own SparseCPU.h, compiled where the sources exist
# ---------------------------------------------------------------------------
_ref = None


def mixed_krylov(method: str, A: Csr, b, x0=None, rel_tol: float = 1e-10,
                 inner_rel_tol: float = 1e-4, max_outer: int = 100, max_inner: int = 0,
                 hist_cap: int = 64):
    """Restarted FP32/FP64 refinement with an FP32 CG or BiCG inner solve (one column of
    gis_solve_multi_mixed) -> (x, outer residual history, inner iterations, status)."""
    m = {"CG": 0, "BiCG": 1}[method]
    x = _c(np.array(np.zeros(A.n) if x0 is None else x0, dtype=np.float64, copy=True),
           np.float64)
    At = transpose(A) if m == 1 else None
    hist = np.zeros(max(hist_cap, 1))
    outer, inner = C.c_int(0), C.c_int(0)
    keep = []

    def vp(a, dt):
        a = _c(a, dt)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    st = lib().or_mixed_krylov(m, A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                               _c(A.val, np.float64),
                               vp(At.rowptr, np.int32) if At else None,
                               vp(At.col, np.int32) if At else None,
                               vp(At.val, np.float64) if At else None,
                               _c(b, np.float64), x, rel_tol, inner_rel_tol, max_outer,
                               max_inner, hist, hist_cap, C.byref(outer), C.byref(inner))
    return x, hist[:min(outer.value, hist_cap)].copy(), inner.value, st


def ref_lib():
    """ctypes handle of oracle/_ref/libref_sparsecpu.so, or None if unbuildable here."""
    global _ref
    if _ref is None:
        try:
            build_ref()
        except Exception:
            return None
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_transpose.restype = None
        R.ref_transpose.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]
        R.ref_jacobi_precondition.restype = C.c_int
        R.ref_jacobi_precondition.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p]
        R.ref_csr_save.restype = C.c_int
        R.ref_csr_save.argtypes = [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        R.ref_csr_read_sizes.restype = C.c_int
        R.ref_csr_read_sizes.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        R.ref_csr_read.restype = C.c_int
        R.ref_csr_read.argtypes = [C.c_char_p, _i32p, _i32p, _f64p]
        R.ref_vec_save.restype = C.c_int
        R.ref_vec_save.argtypes = [C.c_char_p, C.c_int, _f64p]
        R.ref_vec_read.restype = C.c_int
        R.ref_vec_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.c_void_p]
        _ref = R
    return _ref


def ref_transpose(A: Csr) -> Csr:
    R = ref_lib()
    t_rowptr = np.zeros(A.n + 1, np.int32)
    t_col = np.zeros(max(A.nnz, 1), np.int32)
    t_val = np.zeros(max(A.nnz, 1))
    R.ref_transpose(A.n, A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                    _c(A.val, np.float64), t_rowptr, t_col, t_val)
    return Csr(A.n, t_rowptr, t_col[:A.nnz], t_val[:A.nnz])


def ref_csr_read(path: str):
    """The reference's SparseCSR::read (suffix dispatch): (n_rows, n_cols, rowptr, col, val),
    or None when it throws."""
    R = ref_lib()
    nr, nc, nnz = C.c_int(0), C.c_int(0), C.c_int(0)
    if R.ref_csr_read_sizes(path.encode(), C.byref(nr), C.byref(nc), C.byref(nnz)) != 0:
        return None
    rp = np.zeros(nr.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1))
    if R.ref_csr_read(path.encode(), rp, col, val) != 0:
        return None
    return nr.value, nc.value, rp, col[:nnz.value], val[:nnz.value]


def ref_csr_save(path: str, n_rows: int, n_cols: int, rowptr, col, val) -> None:
    ref_lib().ref_csr_save(path.encode(), n_rows, n_cols, _c(rowptr, np.int32),
                           _c(col, np.int32), _c(val, np.float64))


def ref_vec_save(path: str, v) -> None:
    v = _c(v, np.float64)
    ref_lib().ref_vec_save(path.encode(), v.shape[0], v)


def ref_vec_read(path: str, mtx: bool):
    R = ref_lib()
    n = C.c_int(0)
    if R.ref_vec_read(path.encode(), int(mtx), C.byref(n), None) != 0:
        return None
    out = np.zeros(max(n.value, 1))
    R.ref_vec_read(path.encode(), int(mtx), C.byref(n), out.ctypes.data_as(C.c_void_p))
    return out[:n.value]


def ref_jacobi_precondition(A: Csr) -> np.ndarray:
    R = ref_lib()
    out = np.zeros(A.n)
    st = R.ref_jacobi_precondition(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                   _c(A.val, np.float64), out)
    if st != 0:
        raise ValueError("D is not invertible")
    return out


# ---- CPU baseline timings (oracle/cpu_baseline.c; bench.py only) -----------------------
_bl = None


def _baseline_lib():
    global _bl
    if _bl is None:
        build()
        L = C.CDLL(BASELINE_PATH)
        L.bl_threads.restype = C.c_int
        L.bl_threads.argtypes = []
        L.bl_cg_csr_omp.restype = C.c_double
        L.bl_cg_csr_omp.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.bl_pcg_openfoam.restype = C.c_int
        L.bl_pcg_openfoam.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                      _f64p, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.bl_cg_ldu.restype = C.c_double
        L.bl_cg_ldu.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                _f64p, C.c_int]
        L.bl_pcg_openfoam_omp.restype = C.c_int
        L.bl_pcg_openfoam_omp.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                          C.c_int, C.POINTER(C.c_double)]
        L.bl_pbicg_openfoam_omp.restype = C.c_int
        L.bl_pbicg_openfoam_omp.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                            _f64p, _f64p, C.c_double, C.c_int,
                                            C.POINTER(C.c_double)]
        _bl = L
    return _bl


def baseline_threads() -> int:
    return int(_baseline_lib().bl_threads())


def time_cg_csr_omp(A: Csr, b, iters: int):
    """Seconds for `iters` FP64 CG iterations (plus set-up residual) on all OpenMP threads."""
    import time
    x = np.zeros(A.n)
    t0 = time.perf_counter()
    res = _baseline_lib().bl_cg_csr_omp(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                        _c(A.val, np.float64), _c(b, np.float64), x, iters)
    return time.perf_counter() - t0, res


def time_cg_ldu(sys, iters: int):
    """Seconds for `iters` FP64 CG iterations on one thread, matrix applied by the LDU
    face loop."""
    import time
    x = np.zeros(sys.n_cells)
    t0 = time.perf_counter()
    res = _baseline_lib().bl_cg_ldu(sys.n_cells, sys.n_faces, _c(sys.lower_addr, np.int32),
                                    _c(sys.upper_addr, np.int32), _c(sys.diag, np.float64),
                                    _c(sys.lower, np.float64), _c(sys.upper, np.float64),
                                    _c(sys.b, np.float64), x, iters)
    return time.perf_counter() - t0, res


def pcg_openfoam(sys, tolerance: float, max_iter: int):
    """OpenFOAM's PCG (preconditioner none) on one thread: (x, iterations, final scaled
    residual, seconds)."""
    import time
    x = np.zeros(sys.n_cells)
    res = C.c_double(0.0)
    t0 = time.perf_counter()
    it = _baseline_lib().bl_pcg_openfoam(sys.n_cells, sys.n_faces, _c(sys.lower_addr, np.int32),
                                         _c(sys.upper_addr, np.int32), _c(sys.diag, np.float64),
                                         _c(sys.lower, np.float64), _c(sys.upper, np.float64),
                                         _c(sys.b, np.float64), x, tolerance, max_iter,
                                         C.byref(res))
    return x, int(it), res.value, time.perf_counter() - t0


def pcg_openfoam_omp(A: Csr, b, tolerance: float, max_iter: int):
    """OpenFOAM's PCG (preconditioner none) on all OpenMP threads, matrix applied row
    by row: (x, iterations, final scaled residual, seconds)."""
    import time
    x = np.zeros(A.n)
    res = C.c_double(0.0)
    t0 = time.perf_counter()
    it = _baseline_lib().bl_pcg_openfoam_omp(A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32),
                                             _c(A.val, np.float64), _c(b, np.float64), x,
                                             tolerance, max_iter, C.byref(res))
    return x, int(it), res.value, time.perf_counter() - t0


def pbicg_openfoam_omp(A: Csr, b, tolerance: float, max_iter: int, At: Csr | None = None):
    """OpenFOAM's PBiCG (preconditioner none) on all OpenMP threads: (x, iterations,
    final scaled residual, seconds)."""
    import time
    if At is None:
        At = transpose(A)
    x = np.zeros(A.n)
    res = C.c_double(0.0)
    t0 = time.perf_counter()
    it = _baseline_lib().bl_pbicg_openfoam_omp(
        A.n, _c(A.rowptr, np.int32), _c(A.col, np.int32), _c(A.val, np.float64),
        _c(At.rowptr, np.int32), _c(At.col, np.int32), _c(At.val, np.float64),
        _c(b, np.float64), x, tolerance, max_iter, C.byref(res))
    return x, int(it), res.value, time.perf_counter() - t0
