"""Argument checks of the Python binding (cufoam_amd/gis.py), on the CPU.

Every vector handed to the C ABI must have the handle's length (nrhs x N for
several right-hand sides), float64 values, contiguous storage, and -- for
device tensors -- the handle's GPU; otherwise the binding raises GisError
(ERR_ARG) before any C call, since the ABI takes bare pointers and would read
or write past a short buffer.  The handle here has no library behind it: a
call that reached the C ABI would fail the test.
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from cufoam_amd.gis import ERR_ARG, ERR_STATE, Gis, GisError


class _NoCalls:
    def __getattr__(self, name):
        raise AssertionError(f"{name} reached the C ABI")


def _handle(n=27, n_faces=54):
    g = object.__new__(Gis)
    g._L = _NoCalls()
    g._h = None
    g.device = 0
    g.n, g.n_faces, g.nnz, g.n_cpl = n, n_faces, n + 2 * n_faces, 0
    return g


def _raises(status, fn, *a, **k):
    with pytest.raises(GisError) as e:
        fn(*a, **k)
    assert e.value.status == status, str(e.value)


def test_short_or_long_host_vectors_are_rejected():
    g = _handle()
    b = np.ones(27)
    for bad in (np.ones(26), np.ones(28), np.ones((3, 9, 2))):
        _raises(ERR_ARG, g.solve, "CG", bad)
        _raises(ERR_ARG, g.solve, "CG", b, bad)
        _raises(ERR_ARG, g.solve_pcg, "none", bad)
        _raises(ERR_ARG, g.solve_mixed, bad)
        _raises(ERR_ARG, g.solve_mixed, b, bad)
        _raises(ERR_ARG, g.norm_factor, bad, b)
        _raises(ERR_ARG, g.norm_factor, b, bad)
    _raises(ERR_ARG, g.set_values, np.ones(26), np.ones(54), np.ones(54))
    _raises(ERR_ARG, g.set_values, np.ones(27), np.ones(53), np.ones(54))
    _raises(ERR_ARG, g.set_values, np.ones(27), np.ones(54), np.ones(55))


def test_multi_rhs_shapes_are_rejected():
    g = _handle()
    for bad in (np.ones((2, 26)), np.ones(27), np.ones((9, 27)), np.ones((0, 27))):
        _raises(ERR_ARG, g.solve_multi, "CG", bad)
        _raises(ERR_ARG, g.solve_multi_mixed, "BiCG", bad)
        if bad.ndim == 2:   # a single loss vector (1-D) is one column
            _raises(ERR_ARG, g.adjoint, bad, np.ones(27))
    _raises(ERR_ARG, g.solve_multi, "CG", np.ones((2, 27)), np.ones((2, 26)))
    _raises(ERR_ARG, g.adjoint, np.ones((2, 27)), np.ones(26))


def test_foam_solve_needs_a_full_float64_psi():
    g = _handle()
    src = np.ones(27)
    for psi in (np.ones(26), np.ones(27, np.float32), np.ones(54)[::2], [0.0] * 27):
        _raises(ERR_ARG, g.foam_solve, "CG", psi, src, 1e-8, 100)
    _raises(ERR_ARG, g.foam_solve, "CG", np.zeros(27), np.ones(26), 1e-8, 100)


def test_batch_vectors_follow_the_batch_sizes():
    g = _handle()
    _raises(ERR_STATE, g.batch_values, np.ones(3), np.ones(0), np.ones(0))
    g.batch_sizes = np.array([8, 27], np.int32)
    g.batch_faces = np.array([12, 54], np.int32)
    _raises(ERR_ARG, g.batch_values, np.ones(34), np.ones(66), np.ones(66))
    _raises(ERR_ARG, g.batch_values, np.ones(35), np.ones(65), np.ones(66))
    _raises(ERR_ARG, g.batch_solve, np.ones(34), None, 1e-8)
    _raises(ERR_ARG, g.batch_solve, np.ones(35), np.ones(36), 1e-8)


def test_pattern_addressing_and_state():
    g = _handle(n=0, n_faces=0)
    _raises(ERR_STATE, g.solve, "CG", np.ones(4))
    _raises(ERR_STATE, g.set_values, np.ones(4), np.ones(0), np.ones(0))
    _raises(ERR_ARG, g.set_pattern, 4, np.zeros(3, np.int32), np.zeros(2, np.int32))


def test_device_only_calls_reject_host_arrays():
    g = _handle()
    b = np.ones(27)
    _raises(ERR_ARG, g.dot, b, b)
    _raises(ERR_ARG, g.spmv, b)
    _raises(ERR_ARG, g.spmv_async, b, b)
    _raises(ERR_ARG, g.norm_factor_noalloc, b, b)


def test_device_tensor_checks_with_cpu_stand_ins(monkeypatch):
    """The device-tensor checks (dtype, contiguity, length, device index) on CPU
    tensors posing as device tensors."""
    torch = pytest.importorskip("torch")
    import cufoam_amd.gis as G
    monkeypatch.setattr(G, "_is_dev", lambda a: isinstance(a, torch.Tensor))
    g = _handle()
    good = torch.ones(27, dtype=torch.float64)
    assert g._dev_vec(good, 27, "x") is good
    _raises(ERR_ARG, g._dev_vec, torch.ones(27, dtype=torch.float32), 27, "x")
    _raises(ERR_ARG, g._dev_vec, torch.ones(26, dtype=torch.float64), 27, "x")
    _raises(ERR_ARG, g._dev_vec, torch.ones(54, dtype=torch.float64)[::2], 27, "x")
    _raises(ERR_ARG, g.solve, "CG", good, np.zeros(27))          # device b, host x0
    _raises(ERR_ARG, g.solve, "CG", torch.ones(27, dtype=torch.float32))
    _raises(ERR_ARG, g.norm_factor_noalloc, good, torch.ones(26, dtype=torch.float64))


def test_system_of_the_right_size_passes_the_checks():
    g = _handle()
    s = mesh.laplacian(3, seed=0)
    x, b, dev = g._xb(s.b, None, "solve")
    assert not dev and x.shape == (27,) and np.all(x == 0) and b.flags.c_contiguous
    X, B, dev = g._xb(np.ones((2, 27)), None, "multi", 54, (2, 27))
    assert X.shape == (2, 27) and B.shape == (2, 27)


def test_binding_has_every_public_method():
    for name in ("set_pattern", "set_values", "set_interfaces", "set_coupled_values",
                 "set_system", "get_csr", "layout", "spmv_bytes", "spmv", "spmv_async", "dot",
                 "norm_factor", "norm_factor_noalloc", "set_schedule", "get_schedule",
                 "autotune", "set_timing", "get_timing", "synchronize", "solve_pcg",
                 "set_slot_fault", "set_gamg", "gamg_levels", "gamg_level", "solve",
                 "solve_mixed", "solve_multi", "solve_multi_mixed", "adjoint",
                 "batch_set_systems", "batch_values", "batch_solve", "foam_solve"):
        assert callable(getattr(Gis, name, None)), name
