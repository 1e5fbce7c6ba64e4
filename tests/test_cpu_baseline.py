"""The CPU baseline's restatements of OpenFOAM's PCG and PBiCG (preconditioner
none) -- timing routines of bench.py, checked here so the figures it reports
are for working solvers:

  - the threaded PCG follows the one-thread face-loop PCG (the same algorithm,
    another matvec order): equal iteration counts up to one, equal residuals
    to rounding;
  - on a symmetric matrix PBiCG's shadow recurrence equals the primal one, so
    it takes PCG's iterations;
  - on a nonsymmetric matrix PBiCG reaches the requested scaled residual, and
    the residual it reports equals sum|b - A x| / normFactor recomputed here
    from x with OpenFOAM's normFactor formula (psi = 0: xRef = 0).
"""
import numpy as np

from cufoam_amd import mesh
from oracle import oracle as O


def _scaled_l1(A, b, x):
    r = b - O.spmv(A, x)
    ax0 = np.zeros(A.n)   # xRef * (A 1) with psi = 0
    nf = np.sum(np.abs(O.spmv(A, np.zeros(A.n)) - ax0) + np.abs(b - ax0)) + 1e-20
    return np.sum(np.abs(r)) / nf


def test_threaded_pcg_follows_the_face_loop_pcg():
    s = mesh.laplacian(20, seed=1, shift=1e-3)
    A = O.ldu2csr(s)
    _, it1, res1, _ = O.pcg_openfoam(s, 1e-8, 2000)
    x, it2, res2, _ = O.pcg_openfoam_omp(A, s.b, 1e-8, 2000)
    assert abs(it1 - it2) <= 1 and res2 < 1e-8
    assert abs(_scaled_l1(A, s.b, x) - res2) <= 1e-6 * res2 + 1e-13


def test_pbicg_equals_pcg_on_a_symmetric_matrix():
    s = mesh.laplacian(16, seed=2, shift=1e-2)
    A = O.ldu2csr(s)
    _, it_cg, _, _ = O.pcg_openfoam_omp(A, s.b, 1e-9, 2000)
    _, it_bi, res, _ = O.pbicg_openfoam_omp(A, s.b, 1e-9, 2000)
    assert abs(it_cg - it_bi) <= 1 and res < 1e-9


def test_pbicg_converges_on_a_nonsymmetric_matrix():
    s = mesh.laplacian(18, seed=3, shift=0.1, symmetric=False)
    A = O.ldu2csr(s)
    x, it, res, _ = O.pbicg_openfoam_omp(A, s.b, 1e-9, 3000)
    assert 0 < it < 3000 and res < 1e-9
    assert abs(_scaled_l1(A, s.b, x) - res) <= 1e-6 * res + 1e-13
