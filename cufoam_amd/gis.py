"""ctypes binding of the gis C ABI (include/gis.h, libgis_amd.so).

Host arrays are numpy arrays; device arrays are torch tensors on a ROCm
device (``tensor.is_cuda``) and take the ``_device`` entry points, so data
stays resident in HBM.  Every call raises ``GisError`` on an error status;
solve outcomes (breakdown, iteration cap, NaN) are returned, not raised.
There is no CPU fallback: without the library or a GPU these calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import build as _build

OK, ERR_ARG, ERR_HIP, ERR_SHAPE, ERR_STATE, ERR_NOMEM = 0, 1, 2, 3, 4, 5
DETAILS = {"error": 0, "time": 1}   # what a solve's history records (DetailType)
BREAKDOWN, NOT_CONVERGED, NAN = 10, 11, 12
ERR_SYNC = 13   # a device reduction did not complete: an error, raised

METHODS = {"CG": 0, "BiCG": 1, "Jacobi": 2, "MR": 3, "MRJacobi": 4, "JMR": 5, "MRS": 6,
           "Richardson": 7}
PRECONDITIONERS = {"none": 0, "diagonal": 1, "GAMG": 2, "GAMG32": 3, "DIC": 4, "DILU": 5}


class GisError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gis status {status}: {msg}")
        self.status = status


class _Params(C.Structure):
    _fields_ = [("error", C.c_double), ("max_steps", C.c_longlong), ("omega", C.c_double),
                ("power_steps", C.c_int), ("int_bound", C.c_double), ("detail_type", C.c_int)]


class _Schedule(C.Structure):
    _fields_ = [("spmv64", C.c_int), ("spmv32", C.c_int), ("block_iters", C.c_int),
                ("fp64_direction", C.c_int), ("fp32_direction", C.c_int),
                ("coarsen64", C.c_int), ("coarsen32", C.c_int)]


class _Layout(C.Structure):
    _fields_ = [(f, C.c_longlong) for f in ("n_rows", "nnz", "slices", "sell_slots",
                                            "table_slices", "table_entries",
                                            "byte_mask_slices", "untabled_slots")]


class _Perf(C.Structure):
    _fields_ = [("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("n_iterations", C.c_int), ("status", C.c_int), ("norm_factor", C.c_double)]


_lib = None

# (name, restype, argtypes) of every extern "C" symbol in include/gis.h
_vp, _i, _ll, _d = C.c_void_p, C.c_int, C.c_longlong, C.c_double
_ip, _llp, _dp = C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_double)
SIGNATURES = {
    "gis_default_params": (None, [C.POINTER(_Params)]),
    "gis_create": (_i, [C.POINTER(_vp), _i]),
    "gis_destroy": (_i, [_vp]),
    "gis_set_stream": (_i, [_vp, _vp]),
    "gis_reset_stream": (_i, [_vp]),
    "gis_synchronize": (_i, [_vp]),
    "gis_status_string": (C.c_char_p, [_i]),
    "gis_last_error": (C.c_char_p, [_vp]),
    "gis_set_pattern": (_i, [_vp, _i, _i, _vp, _vp]),
    "gis_set_pattern_device": (_i, [_vp, _i, _i, _vp, _vp]),
    "gis_set_values": (_i, [_vp, _vp, _vp, _vp]),
    "gis_set_values_device": (_i, [_vp, _vp, _vp, _vp]),
    "gis_set_interfaces": (_i, [_vp, _i, _vp, _vp]),
    "gis_set_coupled_values": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gis_set_coupled_values_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gis_get_size": (_i, [_vp, _ip, _llp]),
    "gis_get_layout": (_i, [_vp, C.POINTER(_Layout)]),
    "gis_get_csr": (_i, [_vp, _vp, _vp, _vp]),
    "gis_get_csr_transpose": (_i, [_vp, _vp, _vp, _vp]),
    "gis_spmv_device": (_i, [_vp, _vp, _vp]),
    "gis_spmv_transpose_device": (_i, [_vp, _vp, _vp]),
    "gis_dot_device": (_i, [_vp, _vp, _vp, _dp]),
    "gis_spmv_device_async": (_i, [_vp, _vp, _vp]),
    "gis_spmv32_device_async": (_i, [_vp, _vp, _vp]),
    "gis_set_schedule": (_i, [_vp, C.POINTER(_Schedule)]),
    "gis_get_schedule": (_i, [_vp, C.POINTER(_Schedule)]),
    "gis_autotune": (_i, [_vp, _i, _i, C.POINTER(_Schedule), _ip, _dp]),
    "gis_autotune_clear": (_i, []),
    "gis_set_timing": (_i, [_vp, _i]),
    "gis_get_timing": (_i, [_vp, _i, _dp, _llp]),
    "gis_norm_factor": (_i, [_vp, _vp, _vp, _dp]),
    "gis_norm_factor_device": (_i, [_vp, _vp, _vp, _dp]),
    "gis_norm_factor_device_noalloc": (_i, [_vp, _vp, _vp, _dp]),
    "gis_solve": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_solve_device": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_foam_solve": (_i, [_vp, _i, _vp, _vp, _d, _ll, C.POINTER(_Perf)]),
    "gis_foam_solve_controls": (_i, [_vp, _i, _vp, _vp, _d, _d, _ll, _ll, C.POINTER(_Perf)]),
    "gis_solve_pcg": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_solve_pcg_device": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_foam_solve_pcg": (_i, [_vp, _i, _vp, _vp, _d, _d, _ll, _ll, C.POINTER(_Perf)]),
    "gis_solve_pbicg": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_solve_pbicg_device": (_i, [_vp, _i, _vp, _vp, C.POINTER(_Params), _vp, _i, _ip]),
    "gis_foam_solve_pbicg": (_i, [_vp, _i, _vp, _vp, _d, _d, _ll, _ll, C.POINTER(_Perf)]),
    "gis_set_gamg": (_i, [_vp, _d, _i, _d]),
    "gis_set_slot_fault": (_i, [_vp, _i]),
    "gis_set_timing_sample": (_i, [_vp, _i]),
    "gis_set_symmetric_storage": (_i, [_vp, _i]),
    "gis_get_symmetric_storage": (_i, [_vp, _ip, C.POINTER(C.c_longlong)]),
    "gis_gamg_levels": (_i, [_vp, _ip]),
    "gis_gamg_level": (_i, [_vp, _i, _ip, _ip, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gis_solve_multi": (_i, [_vp, _i, _i, _vp, _vp, C.POINTER(_Params), _i, _vp, _vp, _vp]),
    "gis_solve_multi_device": (_i, [_vp, _i, _i, _vp, _vp, C.POINTER(_Params), _i, _vp, _vp,
                                    _vp]),
    "gis_solve_multi_mixed": (_i, [_vp, _i, _i, _vp, _vp, _d, _d, _ll, _ll, _vp, _i, _vp, _vp,
                                   _vp, _vp]),
    "gis_solve_multi_mixed_device": (_i, [_vp, _i, _i, _vp, _vp, _d, _d, _ll, _ll, _vp, _i,
                                          _vp, _vp, _vp, _vp]),
    "gis_adjoint": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_Params), _vp,
                         _vp]),
    "gis_adjoint_device": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                                C.POINTER(_Params), _vp, _vp]),
    "gis_batch_set_pattern": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "gis_batch_set_values": (_i, [_vp, _vp, _vp, _vp]),
    "gis_batch_set_values_device": (_i, [_vp, _vp, _vp, _vp]),
    "gis_batch_solve": (_i, [_vp, _vp, _vp, _vp, _ll, _i, _vp, _vp, _vp]),
    "gis_batch_solve_device": (_i, [_vp, _vp, _vp, _vp, _ll, _i, _vp, _vp, _vp]),
    "gis_solve_mixed": (_i, [_vp, _vp, _vp, _d, _d, _ll, _i, _vp, _i, _ip, _ip, _dp]),
    "gis_solve_mixed_device": (_i, [_vp, _vp, _vp, _d, _d, _ll, _i, _vp, _i, _ip, _ip, _dp]),
}


def library_path() -> str:
    return _build.LIB


def load(build_if_missing: bool = True):
    """Load libgis_amd.so (building it first only if it does not exist)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        if not build_if_missing:
            raise GisError(ERR_STATE, f"{path} is not built")
        _build.build()
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    _lib = L
    return L


def _is_dev(a) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def _host(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    if _is_dev(a):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Performance:
    """solverPerformance as cuFoam's wrappers fill it (cuFoamCG.C:80-83)."""
    solver: str
    initial_residual: float
    final_residual: float
    n_iterations: int
    status: int
    norm_factor: float


class Gis:
    """One handle: a device, a stream, a cached LDU pattern and its values."""

    def __init__(self, device: int = 0):
        self._L = load()
        self._h = C.c_void_p()
        st = self._L.gis_create(C.byref(self._h), device)
        if st != OK:
            raise GisError(st, "gis_create failed (is a GPU visible?)")
        self.device = device
        self.n = 0
        self.n_faces = 0
        self.nnz = 0
        self.use_torch_stream()

    def use_torch_stream(self):
        """Run on torch's current stream of this device, so device tensors that torch
        produced are complete before gis reads them (the handle's own stream is
        non-blocking and would not wait for them)."""
        try:
            import torch
        except ImportError:
            return
        if torch.cuda.is_available():
            self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h:
            self._L.gis_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- argument checks (every vector handed to the C ABI) ----------------------
    def _dev_vec(self, a, count: int, what: str, itemsize: int = 8):
        """A device tensor the kernels may read or write as `count` contiguous
        float64 (itemsize 8) or float32 (itemsize 4) entries on this handle's GPU."""
        import torch
        want = torch.float64 if itemsize == 8 else torch.float32
        if a.dtype != want:
            raise GisError(ERR_ARG, f"{what}: needs {want}, got {a.dtype}")
        if not a.is_contiguous():
            raise GisError(ERR_ARG, f"{what}: tensor is not contiguous")
        if a.numel() != count:
            raise GisError(ERR_ARG, f"{what}: needs {count} entries, got {a.numel()}")
        if a.device.index is not None and a.device.index != self.device:
            raise GisError(ERR_ARG, f"{what}: tensor is on {a.device}, handle on cuda:{self.device}")
        return a

    def _host_vec(self, a, count: int, what: str, copy: bool = False, dtype=np.float64):
        """A host array of `count` entries as contiguous `dtype` (a fresh copy when
        `copy`: the C ABI writes into it)."""
        if _is_dev(a):
            raise GisError(ERR_ARG, f"{what}: mixes a device tensor with host arrays")
        h = np.array(a, dtype=dtype, copy=True) if copy else np.ascontiguousarray(a, dtype=dtype)
        h = np.ascontiguousarray(h)
        if h.size != count:
            raise GisError(ERR_ARG, f"{what}: needs {count} entries, got {h.size}")
        return h

    def _vecs(self, count: int, what: str, **arrays):
        """Checks a group of vectors that go to one call: all device tensors (checked
        in place) or all host arrays (converted); names ending in '_out' are copied."""
        dev = [_is_dev(a) for a in arrays.values() if a is not None]
        if any(dev) and not all(dev):
            raise GisError(ERR_ARG, f"{what}: mixes device tensors and host arrays")
        out = {}
        for name, a in arrays.items():
            if a is None:
                out[name] = None
            elif _is_dev(a):
                out[name] = self._dev_vec(a, count, f"{what} {name}")
            else:
                out[name] = self._host_vec(a, count, f"{what} {name}", copy=name.endswith("_out"))
        return out

    def _xb(self, b, x0, what: str, count: int | None = None, shape=None):
        """(x, b, on_device) for a solve: device tensors are checked and x0 is
        updated in place (zeros when None); host arrays are converted and x is a
        fresh copy of x0."""
        self._need_pattern(what)
        count = self.n if count is None else count
        if _is_dev(b):
            import torch
            x = x0 if x0 is not None else torch.zeros_like(b)
            v = self._vecs(count, what, b=b, x=x)
            return v["x"], v["b"], True
        if x0 is None:
            x0 = np.zeros(count)
        v = self._vecs(count, what, b=b, x_out=x0)
        x, bb = v["x_out"], v["b"]
        if shape is not None:
            x, bb = x.reshape(shape), bb.reshape(shape)
        return x, bb, False

    def _nrhs(self, B, what: str) -> int:
        """Columns of a right-hand-side block of shape (nrhs, N)."""
        self._need_pattern(what)
        shape = tuple(B.shape) if hasattr(B, "shape") else np.shape(B)
        if len(shape) != 2 or shape[1] != self.n or not 1 <= shape[0] <= 8:
            raise GisError(ERR_ARG, f"{what}: needs shape (nrhs, {self.n}) with 1 <= nrhs <= 8, "
                                    f"got {shape}")
        return int(shape[0])

    def _need_pattern(self, what: str):
        if self.n <= 0:
            raise GisError(ERR_STATE, f"{what}: set_pattern first (vector length unknown)")

    def _check(self, st: int, what: str) -> int:
        if st != OK and (st < BREAKDOWN or st == ERR_SYNC):
            msg = self._L.gis_last_error(self._h)
            raise GisError(st, f"{what}: {msg.decode() if msg else ''}")
        return st

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: int | None):
        """Run on the given hipStream_t (an int; 0 is the default stream); None goes
        back to the handle's own stream."""
        if stream_ptr is None:
            self._check(self._L.gis_reset_stream(self._h), "reset_stream")
        else:
            self._check(self._L.gis_set_stream(self._h, C.c_void_p(stream_ptr)), "set_stream")

    # -- matrix --------------------------------------------------------------
    def set_pattern(self, n_cells: int, lower_addr, upper_addr):
        if _is_dev(lower_addr) != _is_dev(upper_addr):
            raise GisError(ERR_ARG, "set_pattern: mixes device tensors and host arrays")
        if _is_dev(lower_addr):
            import torch
            m = int(lower_addr.numel())
            for name, a in (("lower_addr", lower_addr), ("upper_addr", upper_addr)):
                if a.dtype != torch.int32 or not a.is_contiguous() or a.numel() != m:
                    raise GisError(ERR_ARG, f"set_pattern: {name} must be contiguous int32 "
                                            f"of {m} entries")
            st = self._L.gis_set_pattern_device(self._h, n_cells, m, _ptr(lower_addr),
                                                _ptr(upper_addr))
        else:
            la = _host(lower_addr, np.int32).ravel()
            ua = _host(upper_addr, np.int32).ravel()
            if la.shape != ua.shape:
                raise GisError(ERR_ARG, "set_pattern: lower_addr and upper_addr differ in length")
            m = la.shape[0]
            st = self._L.gis_set_pattern(self._h, n_cells, m, _ptr(la), _ptr(ua))
        self._check(st, "set_pattern")
        self.n_faces = m
        self.n_cpl = 0
        n = C.c_int(0)
        nnz = C.c_longlong(0)
        self._check(self._L.gis_get_size(self._h, C.byref(n), C.byref(nnz)), "get_size")
        self.n, self.nnz = n.value, nnz.value

    def set_values(self, diag, lower, upper):
        self._need_pattern("set_values")
        d = self._vecs(self.n, "set_values", diag=diag)["diag"]
        f = self._vecs(self.n_faces, "set_values", lower=lower, upper=upper)
        if _is_dev(d) != _is_dev(f["lower"]):
            raise GisError(ERR_ARG, "set_values: mixes device tensors and host arrays")
        if _is_dev(d):
            st = self._L.gis_set_values_device(self._h, _ptr(d), _ptr(f["lower"]),
                                               _ptr(f["upper"]))
        else:
            st = self._L.gis_set_values(self._h, _ptr(d), _ptr(f["lower"]), _ptr(f["upper"]))
        self._check(st, "set_values")

    def set_interfaces(self, row, col):
        """Coupled-interface entries A(row[j], col[j]) (structurally symmetric)."""
        r = _host(row, np.int32)
        c = _host(col, np.int32)
        if r.shape != c.shape:
            raise ValueError("row and col must have the same length")
        self._check(self._L.gis_set_interfaces(self._h, r.shape[0], _ptr(r), _ptr(c)),
                    "set_interfaces")
        self.n_cpl = int(r.shape[0])
        n = C.c_int(0)
        nnz = C.c_longlong(0)
        self._check(self._L.gis_get_size(self._h, C.byref(n), C.byref(nnz)), "get_size")
        self.n, self.nnz = n.value, nnz.value

    def set_coupled_values(self, diag, lower, upper, coeffs):
        self._need_pattern("set_coupled_values")
        d = self._vecs(self.n, "set_coupled_values", diag=diag)["diag"]
        f = self._vecs(self.n_faces, "set_coupled_values", lower=lower, upper=upper)
        cv = self._vecs(getattr(self, "n_cpl", 0), "set_coupled_values", coeffs=coeffs)["coeffs"]
        devs = {_is_dev(a) for a in (d, f["lower"], cv) if a is not None}
        if len(devs) > 1:
            raise GisError(ERR_ARG, "set_coupled_values: mixes device tensors and host arrays")
        fn = (self._L.gis_set_coupled_values_device if _is_dev(d)
              else self._L.gis_set_coupled_values)
        st = fn(self._h, _ptr(d), _ptr(f["lower"]), _ptr(f["upper"]),
                None if cv is None else _ptr(cv))
        self._check(st, "set_coupled_values")

    def set_system(self, sys):
        self.set_pattern(sys.n_cells, sys.lower_addr, sys.upper_addr)
        if getattr(sys, "cpl_row", None) is not None and len(sys.cpl_row):
            self.set_interfaces(sys.cpl_row, sys.cpl_col)
            self.set_coupled_values(sys.diag, sys.lower, sys.upper, sys.cpl_val)
        else:
            self.set_values(sys.diag, sys.lower, sys.upper)

    def layout(self) -> dict:
        """The sliced layouts of the pattern (gis_layout): slice, slot and table counts."""
        lay = _Layout()
        self._check(self._L.gis_get_layout(self._h, C.byref(lay)), "get_layout")
        return {f: getattr(lay, f) for f, _ in _Layout._fields_}

    def set_symmetric_storage(self, enable: bool = True):
        """Switch the symmetric storage of the offset-table schedule on or off
        (gis_set_symmetric_storage; on by default, used where it applies)."""
        self._check(self._L.gis_set_symmetric_storage(self._h, 1 if enable else 0),
                    "set_symmetric_storage")

    def symmetric_storage(self) -> tuple[bool, int]:
        """(in use, value slots it keeps) for the current pattern and values."""
        a = C.c_int(0)
        n = C.c_longlong(0)
        self._check(self._L.gis_get_symmetric_storage(self._h, C.byref(a), C.byref(n)),
                    "get_symmetric_storage")
        return bool(a.value), int(n.value)

    def spmv_bytes(self, value_size: int, vector_size: int, schedule: int = 6) -> int:
        """Bytes one SpMV of this pattern must move in the given layout: the stored
        matrix (values, column indices or offset tables and row masks, slice
        offsets), x read once and y written once.  Schedule 6 (column-offset
        tables) and 0 (sliced rows) are modelled; others count the CSR bytes.
        Under the symmetric storage the FP64 schedule 6 stores only the kept slots,
        plus a 16-byte record per table entry."""
        lay = self.layout()
        n, sv, sx = lay["n_rows"], value_size, vector_size
        vec = 2 * n * sx
        if schedule == 6:
            tabled_rows = min(64 * lay["table_slices"], n)
            byte_rows = min(64 * lay["byte_mask_slices"], tabled_rows)
            sym, kept = self.symmetric_storage()
            vals = kept * sv + 16 * lay["table_entries"] if sym and sv == 8 else \
                64 * lay["table_entries"] * sv
            return (vals + byte_rows + 4 * (tabled_rows - byte_rows)
                    + 4 * (lay["table_entries"] + lay["slices"] + 1)
                    + lay["untabled_slots"] * (sv + 4) + vec)
        if schedule in (0, 4):
            return lay["sell_slots"] * (sv + 4) + 4 * (lay["slices"] + 1) + vec
        return lay["nnz"] * (sv + 4) + 4 * (n + 1) + vec

    def get_csr(self, transpose: bool = False):
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.float64)
        f = self._L.gis_get_csr_transpose if transpose else self._L.gis_get_csr
        self._check(f(self._h, _ptr(rp), _ptr(col), _ptr(val)), "get_csr")
        return rp, col[:self.nnz], val[:self.nnz]

    # -- kernels on device tensors ---------------------------------------------
    def spmv(self, x, y=None, transpose: bool = False):
        import torch
        self._need_pattern("spmv")
        if not _is_dev(x) or (y is not None and not _is_dev(y)):
            raise GisError(ERR_ARG, "spmv: x and y must be device tensors")
        self._dev_vec(x, self.n, "spmv x")
        if y is None:
            y = torch.empty_like(x)
        self._dev_vec(y, self.n, "spmv y")
        f = self._L.gis_spmv_transpose_device if transpose else self._L.gis_spmv_device
        self._check(f(self._h, _ptr(x), _ptr(y)), "spmv")
        return y

    def spmv_async(self, x, y):
        self._need_pattern("spmv_async")
        if not (_is_dev(x) and _is_dev(y)):
            raise GisError(ERR_ARG, "spmv_async: x and y must be device tensors")
        size = 4 if x.dtype.itemsize == 4 else 8
        self._dev_vec(x, self.n, "spmv_async x", size)
        self._dev_vec(y, self.n, "spmv_async y", size)
        f = self._L.gis_spmv32_device_async if size == 4 else self._L.gis_spmv_device_async
        self._check(f(self._h, _ptr(x), _ptr(y)), "spmv_async")

    def dot(self, a, b) -> float:
        """(a, b) of two device vectors (torch CUDA tensors, float64)."""
        self._need_pattern("dot")
        if not (_is_dev(a) and _is_dev(b)):
            raise GisError(ERR_ARG, "dot: a and b must be device tensors")
        self._vecs(self.n, "dot", a=a, b=b)
        out = C.c_double(0.0)
        self._check(self._L.gis_dot_device(self._h, _ptr(a), _ptr(b), C.byref(out)), "dot")
        return out.value

    def norm_factor(self, b, x) -> float:
        """NormFactor(A, b, x) (gis/gis.cu:106-123)."""
        self._need_pattern("norm_factor")
        v = self._vecs(self.n, "norm_factor", b=b, x=x)
        out = C.c_double(0.0)
        f = self._L.gis_norm_factor_device if _is_dev(b) else self._L.gis_norm_factor
        self._check(f(self._h, _ptr(v["b"]), _ptr(v["x"]), C.byref(out)), "norm_factor")
        return out.value

    def norm_factor_noalloc(self, b, x) -> float:
        """NormFactor of device tensors through the entry point that allocates no
        device memory (its workspace is the library's fixed per-GPU array)."""
        self._need_pattern("norm_factor_noalloc")
        if not (_is_dev(b) and _is_dev(x)):
            raise GisError(ERR_ARG, "norm_factor_noalloc: b and x must be device tensors")
        self._vecs(self.n, "norm_factor_noalloc", b=b, x=x)
        out = C.c_double(0.0)
        self._check(self._L.gis_norm_factor_device_noalloc(self._h, _ptr(b), _ptr(x),
                                                           C.byref(out)), "norm_factor_noalloc")
        return out.value

    SCHEDULE_FIELDS = ("spmv64", "spmv32", "block_iters", "fp64_direction", "fp32_direction",
                       "coarsen64", "coarsen32")

    def set_schedule(self, spmv64: int = -1, spmv32: int = -1, block_iters: int = -1,
                     fp64_direction: int = -1, fp32_direction: int = -1, coarsen64: int = -1,
                     coarsen32: int = -1):
        """Tuning only (bit-identical results); negative values keep the current field."""
        s = _Schedule(spmv64, spmv32, block_iters, fp64_direction, fp32_direction, coarsen64,
                      coarsen32)
        self._check(self._L.gis_set_schedule(self._h, C.byref(s)), "set_schedule")

    def get_schedule(self) -> dict:
        s = _Schedule()
        self._check(self._L.gis_get_schedule(self._h, C.byref(s)), "get_schedule")
        return {f: getattr(s, f) for f in self.SCHEDULE_FIELDS}

    def autotune(self, fp64: bool = True, fp32: bool = True, iterations: int = 0):
        """Time every candidate schedule on this matrix, apply and cache the fastest.
        Returns (schedule dict, from_cache, best_ms[fp64, fp32])."""
        s = _Schedule()
        cached = C.c_int(0)
        ms = (C.c_double * 2)()
        targets = (1 if fp64 else 0) | (2 if fp32 else 0)
        self._check(self._L.gis_autotune(self._h, targets, iterations, C.byref(s),
                                         C.byref(cached), ms), "autotune")
        return ({f: getattr(s, f) for f in self.SCHEDULE_FIELDS}, bool(cached.value),
                (ms[0], ms[1]))

    @staticmethod
    def autotune_clear():
        load().gis_autotune_clear()

    TIMER_SLOTS = {"iter_spmv": 0, "iter_update": 1, "fp64_update": 2, "spmv": 3,
                   "iter_direction": 4}

    def set_timing(self, *slots, period: int = 1):
        """Record HIP-event durations of the named slots (none: stop timing), one
        launch in `period` of each."""
        mask = 0
        for s in slots:
            mask |= 1 << (self.TIMER_SLOTS[s] if isinstance(s, str) else int(s))
        self._check(self._L.gis_set_timing_sample(self._h, int(period)), "set_timing")
        self._check(self._L.gis_set_timing(self._h, mask), "set_timing")

    def get_timing(self, slot) -> tuple[float, int]:
        """(total milliseconds, launches) recorded for a timer slot since set_timing."""
        s = self.TIMER_SLOTS[slot] if isinstance(slot, str) else int(slot)
        ms, cnt = C.c_double(0.0), C.c_longlong(0)
        self._check(self._L.gis_get_timing(self._h, s, C.byref(ms), C.byref(cnt)),
                    "get_timing")
        return ms.value, cnt.value

    def synchronize(self):
        self._check(self._L.gis_synchronize(self._h), "synchronize")

    # -- solvers ---------------------------------------------------------------
    def solve_pcg(self, precond: str, b, x0=None, error: float = 1e-4, max_steps: int = 0,
                  hist_cap: int | None = None, detail: str = "error"):
        """Preconditioned CG (precond none / diagonal / GAMG) -> (x, hist, status).
        detail "time": hist holds the seconds since the loop started instead of norms."""
        p = _Params(error, max_steps, 0.0, 30, 1e-10, DETAILS[detail])
        if hist_cap is None:
            hist_cap = (max_steps if max_steps > 0 else 10000) + 1
        hist = np.zeros(max(hist_cap, 1))
        hl = C.c_int(0)
        pc = PRECONDITIONERS[precond]
        x, b, dev = self._xb(b, x0, "solve_pcg")
        f = self._L.gis_solve_pcg_device if dev else self._L.gis_solve_pcg
        st = f(self._h, pc, _ptr(x), _ptr(b), C.byref(p), _ptr(hist), hist_cap, C.byref(hl))
        self._check(st, f"solve_pcg({precond})")
        self.last_iterations = max(hl.value - 1, 0)
        return x, hist[:min(hl.value, hist_cap)].copy(), st

    def solve_pbicg(self, precond: str, b, x0=None, error: float = 1e-4, max_steps: int = 0,
                    int_bound: float = 1e-10, hist_cap: int | None = None,
                    detail: str = "error"):
        """BiCG preconditioned by none / diagonal / DILU -> (x, hist, status).
        Breakdown (status 11) is returned, not raised, as for solve("BiCG")."""
        p = _Params(error, max_steps, 0.0, 30, int_bound, DETAILS[detail])
        if hist_cap is None:
            hist_cap = (max_steps if max_steps > 0 else 10000) + 1
        hist = np.zeros(max(hist_cap, 1))
        hl = C.c_int(0)
        pc = PRECONDITIONERS[precond]
        x, b, dev = self._xb(b, x0, "solve_pbicg")
        f = self._L.gis_solve_pbicg_device if dev else self._L.gis_solve_pbicg
        st = f(self._h, pc, _ptr(x), _ptr(b), C.byref(p), _ptr(hist), hist_cap, C.byref(hl))
        self._check(st, f"solve_pbicg({precond})")
        self.last_iterations = max(hl.value - 1, 0)
        return x, hist[:min(hl.value, hist_cap)].copy(), st

    def set_slot_fault(self, polls: int = 0):
        """Failure-path tests only: polls > 0 makes every device reduction give up
        after that many polls (solves then raise GisError 13); 0 restores."""
        self._check(self._L.gis_set_slot_fault(self._h, int(polls)), "set_slot_fault")

    def set_gamg(self, omega: float = 0.8, nu: int = 1, cscale: float = 1.8):
        self._check(self._L.gis_set_gamg(self._h, omega, nu, cscale), "set_gamg")

    def gamg_levels(self) -> int:
        n = C.c_int(0)
        self._check(self._L.gis_gamg_levels(self._h, C.byref(n)), "gamg_levels")
        return n.value

    def gamg_level(self, level: int) -> dict:
        n, m = C.c_int(0), C.c_int(0)
        self._check(self._L.gis_gamg_level(self._h, level, C.byref(n), C.byref(m), None, None,
                                           None, None, None, None), "gamg_level")
        n, m = n.value, m.value
        out = {"n": n, "m": m, "agg": np.zeros(max(n, 1), np.int32),
               "la": np.zeros(max(m, 1), np.int32), "ua": np.zeros(max(m, 1), np.int32),
               "diag": np.zeros(max(n, 1)), "lower": np.zeros(max(m, 1)),
               "upper": np.zeros(max(m, 1))}
        last = level + 1 == self.gamg_levels()
        self._check(self._L.gis_gamg_level(self._h, level, None, None,
                                           None if last else _ptr(out["agg"]), _ptr(out["la"]),
                                           _ptr(out["ua"]), _ptr(out["diag"]),
                                           _ptr(out["lower"]), _ptr(out["upper"])), "gamg_level")
        for k in ("la", "ua", "lower", "upper"):
            out[k] = out[k][:m]
        out["diag"] = out["diag"][:n]
        out["agg"] = None if last else out["agg"][:n]
        return out

    def solve(self, method: str, b, x0=None, error: float = 1e-4, max_steps: int = 0,
              omega: float = 0.0, power_steps: int = 30, int_bound: float = 1e-10,
              hist_cap: int | None = None, detail: str = "error"):
        """SingleGPU::<method>(x, A, b, error, max_steps, ..., detail, detail_type)
        -> (x, hist, status).

        With numpy inputs x is a new numpy array; with device tensors x0 is
        updated in place and returned.  detail "error" (DetailType::Error): hist
        holds ||r_k||_2; "time" (DetailType::Time): the seconds since the loop
        started, 0.0 first.
        """
        p = _Params(error, max_steps, omega, power_steps, int_bound, DETAILS[detail])
        if hist_cap is None:
            hist_cap = (max_steps if max_steps > 0 else 10000) + 1
        hist = np.zeros(max(hist_cap, 1))
        hl = C.c_int(0)
        m = METHODS[method]
        x, b, dev = self._xb(b, x0, f"solve({method})")
        f = self._L.gis_solve_device if dev else self._L.gis_solve
        st = f(self._h, m, _ptr(x), _ptr(b), C.byref(p), _ptr(hist), hist_cap, C.byref(hl))
        self._check(st, f"solve({method})")
        self.last_iterations = max(hl.value - 1, 0)
        return x, hist[:min(hl.value, hist_cap)].copy(), st

    def solve_mixed(self, b, x0=None, rel_tol: float = 1e-10, inner_rel_tol: float = 1e-4,
                    max_iter: int = 0, mode: int = 0, hist_cap: int = 256):
        """Mixed-precision CG -> (x, true-residual history, inner iterations,
        relative residual, status).  Device tensors are updated in place."""
        hist = np.zeros(max(hist_cap, 1))
        hl, it, rr = C.c_int(0), C.c_int(0), C.c_double(0.0)
        x, b, dev = self._xb(b, x0, "solve_mixed")
        f = self._L.gis_solve_mixed_device if dev else self._L.gis_solve_mixed
        st = f(self._h, _ptr(x), _ptr(b), rel_tol, inner_rel_tol, max_iter, mode, _ptr(hist),
               hist_cap, C.byref(hl), C.byref(it), C.byref(rr))
        self._check(st, "solve_mixed")
        return x, hist[:min(hl.value, hist_cap)].copy(), it.value, rr.value, st

    def solve_multi(self, method: str, B, X0=None, error: float = 1e-4, max_steps: int = 0,
                    int_bound: float = 1e-10, transpose: bool = False):
        """Several right-hand sides (B: nrhs x N) -> (X, iterations, final norms, status)."""
        p = _Params(error, max_steps, 0.0, 30, int_bound)
        nrhs = self._nrhs(B, "solve_multi")
        X, B, dev = self._xb(B, X0, f"solve_multi({method})", nrhs * self.n, (nrhs, self.n))
        f = self._L.gis_solve_multi_device if dev else self._L.gis_solve_multi
        it = np.zeros(nrhs, np.int32)
        nf = np.zeros(nrhs)
        sts = np.zeros(nrhs, np.int32)
        st = f(self._h, METHODS[method], nrhs, _ptr(X), _ptr(B), C.byref(p), int(transpose),
               _ptr(it), _ptr(nf), _ptr(sts))
        self._check(st, f"solve_multi({method})")
        return X, it, nf, sts

    def solve_multi_mixed(self, method: str, B, X0=None, rel_tol: float = 1e-10,
                          inner_rel_tol: float = 1e-4, max_outer: int = 100,
                          max_inner: int = 0, hist_cap: int = 64):
        """FP32/FP64 refinement of several right-hand sides (B: nrhs x N) ->
        (X, outer residual histories, inner iterations, relative residuals, status)."""
        nrhs = self._nrhs(B, "solve_multi_mixed")
        X, B, dev = self._xb(B, X0, f"solve_multi_mixed({method})", nrhs * self.n,
                             (nrhs, self.n))
        f = self._L.gis_solve_multi_mixed_device if dev else self._L.gis_solve_multi_mixed
        hist = np.zeros((nrhs, max(hist_cap, 1)))
        outer = np.zeros(nrhs, np.int32)
        it = np.zeros(nrhs, np.int32)
        rel = np.zeros(nrhs)
        sts = np.zeros(nrhs, np.int32)
        st = f(self._h, METHODS[method], nrhs, _ptr(X), _ptr(B), rel_tol, inner_rel_tol,
               max_outer, max_inner, _ptr(hist), hist_cap, _ptr(outer), _ptr(it), _ptr(rel),
               _ptr(sts))
        self._check(st, f"solve_multi_mixed({method})")
        hists = [hist[k, :min(int(outer[k]), hist_cap)].copy() for k in range(nrhs)]
        return X, hists, it, rel, sts

    def adjoint(self, W, x, method: str = "CG", error: float = 1e-12, max_steps: int = 0):
        """Gradients of L_k = W[k] . x through x = A^-1 b -> (lambda = dL/db, d_diag,
        d_lower, d_upper, iterations).  numpy in -> numpy out; device tensors in ->
        device tensors out."""
        self._need_pattern("adjoint")
        m = self.n_faces
        p = _Params(error, max_steps, 0.0, 30, 1e-10)
        if _is_dev(W) != _is_dev(x):
            raise GisError(ERR_ARG, "adjoint: mixes device tensors and host arrays")
        if _is_dev(W):
            import torch
            if W.dim() == 1:
                W = W.unsqueeze(0)
            W = W.contiguous()
            self._nrhs(W, "adjoint")
            x = self._dev_vec(x.contiguous(), self.n, "adjoint x")
            k, n = W.shape
            lam = torch.zeros_like(W)
            dd = torch.zeros_like(W)
            dlo = torch.zeros((k, max(m, 1)), dtype=W.dtype, device=W.device)
            dup = torch.zeros_like(dlo)
            f = self._L.gis_adjoint_device
            x = x.contiguous()
        else:
            W = _host(W, np.float64)
            if W.ndim == 1:
                W = W[None, :]
            self._nrhs(W, "adjoint")
            x = self._host_vec(x, self.n, "adjoint x")
            k, n = W.shape
            lam = np.zeros((k, n))
            dd = np.zeros((k, n))
            dlo = np.zeros((k, max(m, 1)))
            dup = np.zeros((k, max(m, 1)))
            f = self._L.gis_adjoint
        it = np.zeros(k, np.int32)
        sts = np.zeros(k, np.int32)
        st = f(self._h, METHODS[method], k, _ptr(W), _ptr(x), _ptr(lam), _ptr(dd), _ptr(dlo),
               _ptr(dup), C.byref(p), _ptr(it), _ptr(sts))
        self._check(st, "adjoint")
        return lam, dd, dlo[:, :m], dup[:, :m], it

    # -- batched independent systems -------------------------------------------
    def batch_set_systems(self, systems):
        """Stack LduSystems (local addressing each) into one batch on the handle."""
        n = np.array([s.n_cells for s in systems], np.int32)
        m = np.array([s.n_faces for s in systems], np.int32)
        la = np.ascontiguousarray(np.concatenate([s.lower_addr for s in systems]), np.int32)
        ua = np.ascontiguousarray(np.concatenate([s.upper_addr for s in systems]), np.int32)
        self._check(self._L.gis_batch_set_pattern(self._h, len(systems), _ptr(n), _ptr(m),
                                                  _ptr(la), _ptr(ua)), "batch_set_pattern")
        self.batch_sizes = n
        self.batch_faces = m
        self.batch_values(np.concatenate([s.diag for s in systems]),
                          np.concatenate([s.lower for s in systems]),
                          np.concatenate([s.upper for s in systems]))

    def batch_values(self, diag, lower, upper):
        if getattr(self, "batch_sizes", None) is None:
            raise GisError(ERR_STATE, "batch_values: batch_set_systems first")
        d = self._vecs(int(self.batch_sizes.sum()), "batch_values", diag=diag)["diag"]
        f = self._vecs(int(self.batch_faces.sum()), "batch_values", lower=lower, upper=upper)
        if _is_dev(d) != _is_dev(f["lower"]):
            raise GisError(ERR_ARG, "batch_values: mixes device tensors and host arrays")
        fn = self._L.gis_batch_set_values_device if _is_dev(d) else self._L.gis_batch_set_values
        st = fn(self._h, _ptr(d), _ptr(f["lower"]), _ptr(f["upper"]))
        self._check(st, "batch_set_values")

    def batch_solve(self, b, x0=None, error=None, max_steps: int = 0, use_graph: bool = True):
        """CG on every system of the batch -> (x, iterations, final norms, status);
        b and x are the systems' vectors concatenated."""
        S = len(self.batch_sizes)
        err = np.ascontiguousarray(np.broadcast_to(np.asarray(error, np.float64), (S,)))
        it = np.zeros(S, np.int32)
        nf = np.zeros(S)
        sts = np.zeros(S, np.int32)
        total = int(self.batch_sizes.sum())
        if _is_dev(b):
            import torch
            x = x0 if x0 is not None else torch.zeros_like(b)
            v = self._vecs(total, "batch_solve", b=b, x=x)
            b, x = v["b"], v["x"]
            f = self._L.gis_batch_solve_device
        else:
            v = self._vecs(total, "batch_solve", b=b, x_out=x0 if x0 is not None else np.zeros(total))
            b, x = v["b"], v["x_out"]
            f = self._L.gis_batch_solve
        st = f(self._h, _ptr(x), _ptr(b), _ptr(err), max_steps, int(bool(use_graph)), _ptr(it),
               _ptr(nf), _ptr(sts))
        self._check(st, "batch_solve")
        return x, it, nf, sts

    def foam_solve(self, method: str, psi, source, tolerance: float, max_iter: int,
                   rel_tol: float | None = None, min_iter: int = 0,
                   preconditioner: str | None = None):
        """cuFoam<method>::solve: psi (numpy) is updated in place; returns Performance.

        rel_tol / min_iter (OpenFOAM's relTol / minIter) select
        gis_foam_solve_controls; left at None / 0 the call is gis_foam_solve.
        method "PCG" / "PBiCG" with a preconditioner calls gis_foam_solve_pcg /
        gis_foam_solve_pbicg."""
        perf = _Perf()
        self._need_pattern("foam_solve")
        if not (isinstance(psi, np.ndarray) and psi.dtype == np.float64
                and psi.flags.c_contiguous and psi.size == self.n):
            raise GisError(ERR_ARG, f"foam_solve: psi must be a contiguous float64 numpy array "
                                    f"of {self.n} entries (updated in place)")
        source = self._host_vec(source, self.n, "foam_solve source")
        if method in ("PCG", "PBiCG"):
            f = self._L.gis_foam_solve_pcg if method == "PCG" else self._L.gis_foam_solve_pbicg
            st = f(self._h, PRECONDITIONERS[preconditioner or "none"], _ptr(psi), _ptr(source),
                   tolerance, float(rel_tol or 0.0), int(min_iter), max_iter, C.byref(perf))
        elif rel_tol is None and min_iter == 0:
            st = self._L.gis_foam_solve(self._h, METHODS[method], _ptr(psi), _ptr(source),
                                        tolerance, max_iter, C.byref(perf))
        else:
            st = self._L.gis_foam_solve_controls(self._h, METHODS[method], _ptr(psi),
                                                 _ptr(source), tolerance,
                                                 float(rel_tol or 0.0), int(min_iter),
                                                 max_iter, C.byref(perf))
        self._check(st, f"foam_solve({method})")
        return Performance("cuFoam" + method, perf.initial_residual, perf.final_residual,
                           perf.n_iterations, perf.status, perf.norm_factor)
