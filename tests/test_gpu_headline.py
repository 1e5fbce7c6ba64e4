"""The headline workload against the oracle, at its full size.

The system is the bench's: the 256^3 SPD Laplacian (16,777,216 cells,
117,047,296 nonzeros), off-diagonals -(1+0.5u), shift 1e-4, seed 0, b ~ U[-1,1],
x0 = 0, under the autotuned schedule (SellD sliced rows, the bench's layout).
The oracle runs on all host threads (oracle/gis_oracle_par.c, the same bits as
the serial restatement of SolverSingleGPU.h:366-445 and of the mixed CG):

  - FP64 CG and mixed CG for a fixed 40 iterations: histories and x bit for bit;
  - FP64 CG to ||r|| <= 1e-10 ||b||: all 1842 history entries and x bit for bit;
  - mixed CG (reliable updates, delta 1e-2; and the literal restarted
    refinement with inner relTol 1e-4) to the FP64 target: x, history and
    iteration count bit for bit, the true residual recomputed on the host, and
    the error against the FP64 oracle solve.
"""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def headline(gpu):
    print("[case] headline 256^3: mesh, host CSR, device pattern + tune", flush=True)
    s = mesh.laplacian(256, seed=0, shift=1e-4, symmetric=True)
    A = O.ldu2csr(s)
    gpu.set_system(s)
    sched, _, _ = gpu.autotune(iterations=24)
    print(f"[case] schedule {sched}; oracle threads {O.par_threads()}", flush=True)
    bnorm = float(np.linalg.norm(s.b))
    state = {"s": s, "A": A, "bnorm": bnorm}
    yield state
    gpu.set_system(mesh.laplacian(2, seed=0))   # release the 256^3 buffers' use


def _dev(a):
    import torch
    return torch.as_tensor(a, device="cuda:0")


def test_fp64_cg_40_iterations_bit_exact(gpu, headline):
    s, A = headline["s"], headline["A"]
    ref = O.cg_par(A, s.b, s.x0, 0.0, 40)
    x, hist, st = gpu.solve("CG", _dev(s.b), _dev(s.x0), error=0.0, max_steps=40)
    assert len(hist) == 41 == ref.hist_len
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x.cpu().numpy(), ref.x)


@pytest.mark.parametrize("mode,inner", [(0, 1e-2), (1, 1e-4)])
def test_mixed_cg_40_iterations_bit_exact(gpu, headline, mode, inner):
    s, A = headline["s"], headline["A"]
    rx, rh, rit, rst = O.mixed_cg_par(A, s.b, s.x0, 1e-10, inner, 40, mode)
    x, hist, it, rel, st = gpu.solve_mixed(_dev(s.b), _dev(s.x0), 1e-10, inner, 40, mode)
    assert (it, st) == (rit, rst) == (40, 11)
    np.testing.assert_array_equal(hist, rh)
    np.testing.assert_array_equal(x.cpu().numpy(), rx)


def test_fp64_cg_to_1e10_bit_exact(gpu, headline):
    s, A, bnorm = headline["s"], headline["A"], headline["bnorm"]
    err = 1e-10 * bnorm
    print("[case] FP64 CG to 1e-10 on the host", flush=True)
    ref = O.cg_par(A, s.b, s.x0, err, 0, hist_cap=20001)
    x, hist, st = gpu.solve("CG", _dev(s.b), _dev(s.x0), error=err, max_steps=20000,
                            hist_cap=20001)
    print(f"[case] FP64 CG: {len(hist) - 1} iterations (oracle {ref.iterations})", flush=True)
    assert st == 0 and len(hist) == ref.hist_len
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x.cpu().numpy(), ref.x)
    true_rel = O.residual_norm_par(A, s.b, ref.x) / bnorm
    assert true_rel <= 1.1e-10   # recursive residual at 1e-10; the true one tracks it
    headline["x64"] = ref.x


@pytest.mark.parametrize("mode,inner", [(0, 1e-2), (1, 1e-4)])
def test_mixed_cg_to_1e10_bit_exact_and_true_residual(gpu, headline, mode, inner):
    s, A, bnorm = headline["s"], headline["A"], headline["bnorm"]
    print(f"[case] mixed CG mode {mode} inner {inner} to 1e-10 on the host", flush=True)
    rx, rh, rit, rst = O.mixed_cg_par(A, s.b, s.x0, 1e-10, inner, 100000, mode)
    x, hist, it, rel, st = gpu.solve_mixed(_dev(s.b), _dev(s.x0), 1e-10, inner, 0, mode)
    print(f"[case] mixed: {it} inner iterations, {len(hist) - 1} FP64 updates", flush=True)
    assert (st, rst) == (0, 0) and it == rit
    np.testing.assert_array_equal(hist, rh)
    xg = x.cpu().numpy()
    np.testing.assert_array_equal(xg, rx)
    true_rel = O.residual_norm_par(A, s.b, xg) / bnorm
    assert true_rel <= 1e-10 and true_rel == hist[-1] / bnorm
    x64 = headline.get("x64")
    if x64 is None:
        x64 = O.cg_par(A, s.b, s.x0, 1e-10 * bnorm, 0, hist_cap=1).x
        headline["x64"] = x64
    err = float(np.linalg.norm(xg - x64) / np.linalg.norm(x64))
    print(f"[case] ||x_mixed - x_fp64_oracle|| / ||x_fp64_oracle|| = {err:.3e}", flush=True)
    assert err < 1e-7
