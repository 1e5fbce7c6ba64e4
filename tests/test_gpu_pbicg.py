"""GPU parity of BiCG preconditioned by the diagonal and by DILU
(gis_solve_pbicg, gis_foam_solve_pbicg, the cuFoamPBiCG wrapper; no counterpart
in the reference) against the oracle (oracle/gis_oracle.c pbicg_diag,
oracle/gis_oracle_amg.c or_pbicg): histories, iteration counts and solutions
bit for bit on nonsymmetric structured and unstructured systems under every
SpMV schedule, GIS_PRECOND_NONE = BiCG, DILU on symmetric matrices = PCG-DIC,
breakdown, and the argument checks."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _nonsym(n, seed, shift=0.05):
    print(f"[case] nonsymmetric n={n} seed={seed} shift={shift}", flush=True)
    s = mesh.laplacian(n, seed=seed, shift=shift, symmetric=False)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 2.0, s.n_cells)
    return s


def _check(gpu, s, x0, error, max_steps, A=None):
    A = A if A is not None else O.ldu2csr(s)
    ref = O.pbicg("diagonal", A, s.b, x0, error, max_steps)
    x, hist, st = gpu.solve_pbicg("diagonal", s.b, x0, error=error, max_steps=max_steps,
                                  hist_cap=ref.hist_len + 4)
    assert st == ref.status
    assert len(hist) == ref.hist_len, (len(hist), ref.hist_len)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
    return ref


@pytest.mark.parametrize("n,seed", [(1, 0), (3, 1), (12, 2), (33, 3)])
def test_pbicg_diagonal_bit_exact(gpu, n, seed):
    s = _nonsym(n, seed)
    gpu.set_system(s)
    x0 = np.random.default_rng(seed).uniform(-1, 1, s.n_cells)
    _check(gpu, s, x0, 0.0, 25 if n > 1 else 1)
    ref = _check(gpu, s, s.x0, 1e-10 * np.linalg.norm(s.b), 5000)
    assert ref.status == 0 and ref.hist[-1] < 1e-10 * np.linalg.norm(s.b)


@pytest.mark.parametrize("n_cells,n_faces,seed", [(500, 1800, 1), (5000, 20000, 2)])
def test_pbicg_diagonal_unstructured(gpu, n_cells, n_faces, seed):
    s = mesh.random_ldu(n_cells, n_faces, seed=seed, symmetric=False, shift=0.2)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 3.0, n_cells)
    print(f"[case] random LDU n={n_cells} faces={s.n_faces}", flush=True)
    gpu.set_system(s)
    _check(gpu, s, s.x0, 1e-10 * np.linalg.norm(s.b), 5000)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3, 4, 5, 6])
def test_pbicg_schedules(gpu, schedule):
    s = _nonsym(20, 5, shift=1e-2)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    for direction in (0, 1):
        gpu.set_schedule(schedule, -1, -1, direction, -1, 1)
        _check(gpu, s, s.x0, 0.0, 20, A=A)
    gpu.set_schedule(0, -1, -1, -1, -1, 1)


def test_pbicg_none_is_bicg(gpu):
    s = _nonsym(10, 6)
    gpu.set_system(s)
    x, hist, st = gpu.solve_pbicg("none", s.b, s.x0, error=0.0, max_steps=20)
    x2, hist2, st2 = gpu.solve("BiCG", s.b, s.x0, error=0.0, max_steps=20)
    np.testing.assert_array_equal(hist, hist2)
    np.testing.assert_array_equal(x, x2)


def test_pbicg_symmetric_equals_pcg_diagonal(gpu):
    s = mesh.laplacian(16, seed=7, shift=0.05)
    s.diag = s.diag * np.random.default_rng(7).uniform(1.0, 2.0, s.n_cells)
    gpu.set_system(s)
    x, hist, st = gpu.solve_pbicg("diagonal", s.b, s.x0, error=0.0, max_steps=30)
    x2, hist2, st2 = gpu.solve_pcg("diagonal", s.b, s.x0, error=0.0, max_steps=30)
    np.testing.assert_array_equal(hist, hist2)
    np.testing.assert_array_equal(x, x2)


def test_pbicg_breakdown_is_an_outcome(gpu):
    s = _nonsym(6, 8)
    gpu.set_system(s)
    # a breakdown bound above every |xi| stops the first iteration
    ref = O.pbicg("diagonal", O.ldu2csr(s), s.b, s.x0, 0.0, 10, int_bound=1e30)
    x, hist, st = gpu.solve_pbicg("diagonal", s.b, s.x0, error=0.0, max_steps=10,
                                  int_bound=1e30)
    assert st == ref.status == 10 and len(hist) == ref.hist_len == 1
    np.testing.assert_array_equal(x, ref.x)


def _check_dilu(gpu, s, x0, error, max_steps):
    ref = O.pbicg_ldu("DILU", s, s.b, x0, error, max_steps)
    x, hist, st = gpu.solve_pbicg("DILU", s.b, x0, error=error, max_steps=max_steps,
                                  hist_cap=ref.hist_len + 4)
    assert st == ref.status
    assert len(hist) == ref.hist_len, (len(hist), ref.hist_len)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
    return ref


@pytest.mark.parametrize("n,seed", [(1, 0), (4, 1), (13, 2), (24, 3)])
def test_pbicg_dilu_bit_exact(gpu, n, seed):
    s = _nonsym(n, seed, shift=1e-2)
    gpu.set_system(s)
    x0 = np.random.default_rng(seed).uniform(-1, 1, s.n_cells)
    _check_dilu(gpu, s, x0, 0.0, 12 if n > 1 else 1)
    ref = _check_dilu(gpu, s, s.x0, 1e-10 * np.linalg.norm(s.b), 3000)
    assert ref.status == 0


def test_pbicg_dilu_unstructured_values_refresh_and_schedules(gpu):
    s = mesh.random_ldu(4000, 15000, seed=3, symmetric=False, shift=0.2)
    print(f"[case] DILU random LDU n={s.n_cells} faces={s.n_faces}", flush=True)
    gpu.set_system(s)
    _check_dilu(gpu, s, s.x0, 1e-10 * np.linalg.norm(s.b), 3000)
    # new values: the factor is recomputed
    t = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, 1.5 * s.diag, s.lower,
                       0.5 * s.upper, s.b, s.x0)
    gpu.set_values(t.diag, t.lower, t.upper)
    _check_dilu(gpu, t, t.x0, 0.0, 15)
    for sched in (0, 6):
        gpu.set_schedule(sched, -1, -1, -1, -1, 1)
        _check_dilu(gpu, t, t.x0, 0.0, 10)
    gpu.set_schedule(0, -1, -1, -1, -1, 1)


def test_pbicg_dilu_symmetric_equals_pcg_dic(gpu):
    s = mesh.laplacian(12, seed=4, shift=1e-3)
    gpu.set_system(s)
    x, hist, st = gpu.solve_pbicg("DILU", s.b, s.x0, error=0.0, max_steps=30)
    # DIC after DILU on the same handle: the factor is rebuilt for DIC (same bits here)
    x2, hist2, st2 = gpu.solve_pcg("DIC", s.b, s.x0, error=0.0, max_steps=30)
    np.testing.assert_array_equal(hist, hist2)
    np.testing.assert_array_equal(x, x2)


@pytest.mark.parametrize("pc", ["none", "diagonal", "DILU"])
def test_foam_pbicg_wrapper(gpu, pc):
    from cufoam_amd import foam
    s = _nonsym(16, 10, shift=1e-3)
    m = foam.lduMatrix.from_system(s)
    assert m.asymmetric()
    psi = np.random.default_rng(11).uniform(-1, 1, s.n_cells)
    psi_ref = psi.copy()
    solver = foam.New("cuFoamPBiCG", "U", m, {"tolerance": 1e-9, "maxIter": 1000,
                                              "preconditioner": pc, "relTol": 1e-4,
                                              "convergence": "openfoam"}, gis=gpu)
    perf = solver.solve(psi, s.b)
    ref_x, ref = O.foam_solve("PBiCG", s, psi_ref, s.b, 1e-9, 1000, rel_tol=1e-4,
                              preconditioner=pc)
    assert perf.n_iterations == ref.n_iterations > 0
    assert perf.initial_residual == ref.initial_residual
    assert perf.final_residual == ref.final_residual
    np.testing.assert_array_equal(psi, ref_x)


def test_pbicg_argument_checks(gpu):
    from cufoam_amd import foam
    from cufoam_amd.gis import GisError
    s = _nonsym(4, 1)
    gpu.set_system(s)
    for pc in ("GAMG", "GAMG32", "DIC"):
        with pytest.raises(GisError) as e:
            gpu.solve_pbicg(pc, s.b, s.x0, error=1e-8, max_steps=10)
        assert e.value.status == 1
    with pytest.raises(GisError) as e:   # DILU is for BiCG
        gpu.solve_pcg("DILU", s.b, s.x0, error=1e-8, max_steps=10)
    assert e.value.status == 1
    with pytest.raises(foam.FatalError):
        foam.New("cuFoamPBiCG", "U", foam.lduMatrix.from_system(s),
                 {"preconditioner": "DIC"}, gis=gpu)
    # gis_foam_solve and gis_foam_solve_controls take the reference's eight
    # methods only; the preconditioned ones have their own entries
    import ctypes as C
    from cufoam_amd.gis import _Perf
    psi = np.zeros(s.n_cells)
    perf = _Perf()
    p = psi.ctypes.data_as(C.c_void_p)
    b = np.ascontiguousarray(s.b).ctypes.data_as(C.c_void_p)
    for method in (-1, 8, 9, 13):
        assert gpu._L.gis_foam_solve(gpu._h, method, p, b, 1e-8, 10, C.byref(perf)) == 1
        assert gpu._L.gis_foam_solve_controls(gpu._h, method, p, b, 1e-8, 0.0, 0, 10,
                                              C.byref(perf)) == 1
    assert gpu._L.gis_foam_solve_pbicg(gpu._h, 2, p, b, 1e-8, 0.0, 0, 10, C.byref(perf)) == 1
    assert not psi.any()
