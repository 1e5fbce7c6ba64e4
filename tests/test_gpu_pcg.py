"""GPU parity of preconditioned CG (gis_solve_pcg, gis_foam_solve_pcg) against
the oracle's restatement: no preconditioner (= CG), diagonal, and GAMG (the
device hierarchy level by level, then PCG histories and solutions bit for
bit under every SpMV schedule)."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(n, seed, shift=0.01, symmetric=True):
    print(f"[case] n={n} seed={seed} shift={shift} symmetric={symmetric}", flush=True)
    return mesh.laplacian(n, seed=seed, shift=shift, symmetric=symmetric)


def _check(gpu, pc, s, x0, error, max_steps):
    ref = O.pcg(pc, s, s.b, x0, error, max_steps)
    x, hist, st = gpu.solve_pcg(pc, s.b, x0, error=error, max_steps=max_steps,
                                hist_cap=ref.hist_len + 4)
    assert len(hist) == ref.hist_len, (pc, len(hist), ref.hist_len)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
    return ref


@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (16, 2)])
def test_pcg_none_is_cg(gpu, n, seed):
    s = _case(n, seed)
    gpu.set_system(s)
    x0 = np.random.default_rng(seed).uniform(-1, 1, s.n_cells)
    ref = _check(gpu, "none", s, x0, 0.0, 30)
    x, hist, st = gpu.solve("CG", s.b, x0, error=0.0, max_steps=30)
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)


@pytest.mark.parametrize("n,seed", [(3, 0), (12, 1), (32, 2)])
def test_pcg_diagonal_bit_exact(gpu, n, seed):
    s = _case(n, seed, shift=0.05)
    s.diag = s.diag * np.random.default_rng(seed).uniform(1.0, 2.0, s.n_cells)
    gpu.set_system(s)
    _check(gpu, "diagonal", s, s.x0, 0.0, 40)
    _check(gpu, "diagonal", s, s.x0, 1e-8 * np.linalg.norm(s.b), 5000)


@pytest.mark.parametrize("n,symmetric", [(6, True), (16, True), (20, False), (32, True)])
def test_gamg_hierarchy_matches_the_oracle(gpu, n, symmetric):
    s = _case(n, 3, symmetric=symmetric)
    gpu.set_system(s)
    ref = O.Amg(s)
    assert gpu.gamg_levels() == ref.n_levels
    for l in range(ref.n_levels):
        a, b = gpu.gamg_level(l), ref.level(l)
        assert (a["n"], a["m"]) == (b["n"], b["m"]), l
        for k in ("la", "ua", "diag", "lower", "upper"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"level {l} {k}")
        if b["agg"] is None:
            assert a["agg"] is None
        else:
            np.testing.assert_array_equal(a["agg"], b["agg"], err_msg=f"level {l} agg")


@pytest.mark.parametrize("case", ["random", "renumbered"])
def test_gamg_hierarchy_irregular_graphs(gpu, case):
    """The parallel host set-up on graphs with irregular rows: a random
    unstructured pattern, and a structured mesh under a random cell numbering."""
    if case == "random":
        s = mesh.random_ldu(30000, 110000, seed=8, symmetric=True, shift=0.05)
    else:
        base = mesh.laplacian(28, seed=9, shift=1e-3)
        s = mesh.renumber(base, np.random.default_rng(9).permutation(base.n_cells))
    print(f"[case] GAMG hierarchy on {case} n={s.n_cells} faces={s.n_faces}", flush=True)
    gpu.set_system(s)
    ref = O.Amg(s)
    assert gpu.gamg_levels() == ref.n_levels > 2
    for l in range(ref.n_levels):
        a, b = gpu.gamg_level(l), ref.level(l)
        assert (a["n"], a["m"]) == (b["n"], b["m"]), l
        for k in ("la", "ua", "diag", "lower", "upper"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"level {l} {k}")
        if b["agg"] is not None:
            np.testing.assert_array_equal(a["agg"], b["agg"], err_msg=f"level {l} agg")
    _check(gpu, "GAMG", s, s.x0, 0.0, 8)


@pytest.mark.parametrize("pc", ["GAMG", "GAMG32"])
@pytest.mark.parametrize("n,seed", [(4, 0), (8, 1), (16, 2), (24, 3), (40, 4)])
def test_pcg_gamg_bit_exact(gpu, pc, n, seed):
    s = _case(n, seed, shift=1e-3)
    gpu.set_system(s)
    _check(gpu, pc, s, s.x0, 0.0, 12)
    ref = _check(gpu, pc, s, s.x0, 1e-10 * np.linalg.norm(s.b), 1000)
    assert ref.hist[-1] <= 1e-10 * np.linalg.norm(s.b)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("coarsen", [1, 2, 4])
def test_pcg_gamg_schedules(gpu, schedule, coarsen):
    s = _case(20, 5, shift=1e-3)
    gpu.set_system(s)
    gpu.set_schedule(schedule, -1, -1, -1, -1, coarsen)
    _check(gpu, "GAMG", s, s.x0, 0.0, 10)
    _check(gpu, "GAMG32", s, s.x0, 0.0, 10)
    _check(gpu, "diagonal", s, s.x0, 0.0, 10)
    gpu.set_schedule(0, -1, -1, -1, -1, 1)


@pytest.mark.parametrize("omega,nu,cscale", [(0.6, 2, 1.0), (0.8, 3, 1.5)])
def test_pcg_gamg_settings(gpu, omega, nu, cscale):
    s = _case(16, 6, shift=1e-3)
    gpu.set_system(s)
    gpu.set_gamg(omega, nu, cscale)
    try:
        for pc in ("GAMG", "GAMG32"):
            ref = O.pcg(pc, s, s.b, s.x0, 0.0, 10, omega=omega, nu=nu, cscale=cscale)
            x, hist, st = gpu.solve_pcg(pc, s.b, s.x0, error=0.0, max_steps=10)
            np.testing.assert_array_equal(hist, ref.hist)
            np.testing.assert_array_equal(x, ref.x)
    finally:
        gpu.set_gamg()


def test_gamg_values_refresh_keeps_the_agglomeration(gpu):
    s = _case(16, 7, shift=1e-3)
    gpu.set_system(s)
    _check(gpu, "GAMG", s, s.x0, 0.0, 8)
    # doubling every value keeps every weight comparison, hence the aggregates;
    # the coarse values must follow the new values
    t = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, 2.0 * s.diag, 2.0 * s.lower,
                       2.0 * s.upper, s.b, s.x0)
    _check(gpu, "GAMG32", s, s.x0, 0.0, 8)
    gpu.set_values(t.diag, t.lower, t.upper)
    _check(gpu, "GAMG", t, t.x0, 0.0, 8)
    _check(gpu, "GAMG32", t, t.x0, 0.0, 8)


def test_gamg_single_level_and_no_faces(gpu):
    s = _case(4, 9, shift=0.1)
    gpu.set_system(s)
    assert gpu.gamg_levels() == 1
    ref = _check(gpu, "GAMG", s, s.x0, 1e-9 * np.linalg.norm(s.b), 50)
    assert ref.iterations == 1
    _check(gpu, "GAMG32", s, s.x0, 0.0, 6)
    n = 300
    d = mesh.LduSystem(n, np.zeros(0, np.int32), np.zeros(0, np.int32),
                       np.random.default_rng(3).uniform(1, 2, n), np.zeros(0), np.zeros(0),
                       np.random.default_rng(4).uniform(-1, 1, n), np.zeros(n))
    gpu.set_system(d)
    assert gpu.gamg_levels() == 1
    _check(gpu, "GAMG", d, d.x0, 1e-12, 50)
    _check(gpu, "GAMG32", d, d.x0, 0.0, 5)


@pytest.mark.parametrize("pc", ["none", "diagonal", "GAMG", "GAMG32"])
def test_foam_pcg_wrapper(gpu, pc):
    from cufoam_amd import foam
    s = _case(16, 10, shift=1e-3)
    m = foam.lduMatrix.from_system(s)
    psi = np.random.default_rng(11).uniform(-1, 1, s.n_cells)
    psi_ref = psi.copy()
    solver = foam.New("cuFoamPCG", "p", m, {"tolerance": 1e-9, "maxIter": 1000,
                                            "preconditioner": pc, "relTol": 1e-3,
                                            "convergence": "openfoam"}, gis=gpu)
    perf = solver.solve(psi, s.b)
    ref_x, ref = O.foam_solve("PCG", s, psi_ref, s.b, 1e-9, 1000, rel_tol=1e-3,
                              preconditioner=pc)
    assert perf.n_iterations == ref.n_iterations
    assert perf.final_residual == ref.final_residual
    np.testing.assert_array_equal(psi, ref_x)


def test_pcg_argument_checks(gpu):
    from cufoam_amd.gis import GisError
    s = mesh.periodic_laplacian(4, seed=1, shift=0.1)
    gpu.set_system(s)
    with pytest.raises(GisError):
        gpu.solve_pcg("GAMG", s.b, s.x0, error=1e-8, max_steps=10)
    x, hist, st = gpu.solve_pcg("diagonal", s.b, s.x0, error=0.0, max_steps=5)   # allowed
    with pytest.raises(GisError):
        gpu.set_gamg(0.0, 1, 1.0)
    with pytest.raises(GisError):
        gpu.set_gamg(0.8, 0, 1.0)
    with pytest.raises(GisError):
        gpu.set_gamg(0.8, 1, 2.5)
    assert len(hist) == 6


def _sym_random(n_cells, n_faces, seed):
    r = mesh.random_ldu(n_cells, n_faces, seed=seed, symmetric=True)
    print(f"[case] random symmetric LDU n={n_cells} faces={r.n_faces} seed={seed}", flush=True)
    return r


@pytest.mark.parametrize("case", ["hex1", "hex5", "hex16", "hex33", "random", "faceless"])
def test_pcg_dic_bit_exact(gpu, case):
    """DIC sweeps run level by level on the device; every history entry and x equal
    the oracle's face-loop restatement of OpenFOAM's DIC, at fixed iteration counts
    and to convergence (structured, unstructured and face-less systems)."""
    if case.startswith("hex"):
        s = _case(int(case[3:]), 6, shift=1e-3)
    elif case == "random":
        s = _sym_random(4000, 14000, 7)
    else:
        s = mesh.LduSystem(70, np.zeros(0, np.int32), np.zeros(0, np.int32),
                           np.linspace(1.0, 3.0, 70), np.zeros(0), np.zeros(0),
                           np.linspace(-1.0, 1.0, 70), np.zeros(70))
    gpu.set_system(s)
    _check(gpu, "DIC", s, s.x0, 0.0, 15)
    ref = _check(gpu, "DIC", s, s.x0, 1e-10 * np.linalg.norm(s.b), 3000)
    assert ref.hist[-1] <= 1e-10 * np.linalg.norm(s.b)


@pytest.mark.parametrize("schedule", [0, 6])
def test_pcg_dic_after_new_values_and_schedules(gpu, schedule):
    """The DIC factor follows gis_set_values; the SpMV schedule does not change a bit."""
    s = _case(18, 8, shift=1e-2)
    gpu.set_system(s)
    gpu.set_schedule(spmv64=schedule)
    _check(gpu, "DIC", s, s.x0, 0.0, 20)
    s.diag = s.diag * 1.5
    gpu.set_values(s.diag, s.lower, s.upper)
    gpu.set_schedule(spmv64=schedule)
    _check(gpu, "DIC", s, s.x0, 1e-9 * np.linalg.norm(s.b), 2000)


def test_pcg_dic_rejects_interfaces(gpu):
    from cufoam_amd.gis import GisError
    s = mesh.periodic_laplacian(6, seed=1)
    gpu.set_system(s)
    with pytest.raises(GisError) as e:
        gpu.solve_pcg("DIC", s.b, s.x0, error=1e-8, max_steps=10)
    assert e.value.status == 4
