"""GPU parity of coupled interfaces (gis_set_interfaces / gis_set_coupled_values)
against the oracle's coupled CSR: the device CSR and its transpose, SpMV under
every schedule, the eight methods, mixed precision, several right-hand sides,
the OpenFOAM wrapper mirror, and the argument checks."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

METHODS = ["CG", "BiCG", "Jacobi", "MR", "MRJacobi", "JMR", "MRS", "Richardson"]


def _torch():
    import torch
    return torch


def _case(n, seed, symmetric, shift=0.1):
    print(f"[case] periodic n={n} seed={seed} symmetric={symmetric} shift={shift}", flush=True)
    return mesh.periodic_laplacian(n, seed=seed, shift=shift, symmetric=symmetric)


@pytest.mark.parametrize("n", [2, 5, 16])
def test_device_csr_and_transpose_match_the_oracle(gpu, n):
    s = _case(n, 1, False)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    assert gpu.nnz == A.nnz == s.nnz
    rp, col, val = gpu.get_csr()
    np.testing.assert_array_equal(rp, A.rowptr)
    np.testing.assert_array_equal(col, A.col)
    np.testing.assert_array_equal(val, A.val)
    T = O.mirror_transpose(A)
    rp, col, val = gpu.get_csr(transpose=True)
    np.testing.assert_array_equal(rp, T.rowptr)
    np.testing.assert_array_equal(col, T.col)
    np.testing.assert_array_equal(val, T.val)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3, 4, 5, 6])
def test_spmv_schedules_with_couplings(gpu, schedule):
    torch = _torch()
    s = _case(24, 2, False)
    gpu.set_system(s)
    gpu.set_schedule(schedule, schedule, -1)
    A = O.ldu2csr(s)
    x = np.random.default_rng(3).standard_normal(s.n_cells)
    dx = torch.from_numpy(x).cuda()
    assert np.array_equal(gpu.spmv(dx).cpu().numpy(), O.spmv(A, x))
    assert np.array_equal(gpu.spmv(dx, transpose=True).cpu().numpy(),
                          O.spmv(O.mirror_transpose(A), x))
    x32 = dx.float()
    y32 = torch.empty_like(x32)
    gpu.spmv_async(x32, y32)
    gpu.synchronize()
    assert np.array_equal(y32.cpu().numpy(), O.spmv_f32(A, x.astype(np.float32)))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n,seed", [(2, 0), (8, 1), (16, 2)])
def test_methods_bit_exact_with_couplings(gpu, method, n, seed):
    s = _case(n, seed, method in ("CG", "Richardson"))
    gpu.set_system(s)
    A = O.ldu2csr(s)
    At = O.mirror_transpose(A) if method == "BiCG" else None
    x0 = np.random.default_rng(seed).uniform(-1, 1, s.n_cells)
    ref = O.solve(method, A, s.b, x0, 0.0, 30, At=At)
    x, hist, st = gpu.solve(method, s.b, x0, error=0.0, max_steps=30, hist_cap=ref.hist_len + 4)
    assert st == ref.status
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)


@pytest.mark.parametrize("method", ["CG", "BiCG", "MRJacobi"])
def test_coupled_solution_is_the_periodic_solution(gpu, method):
    """Against the same periodic matrix held as ordinary faces (dense solve)."""
    s = _case(12, 3, method == "CG", shift=0.5)
    gpu.set_system(s)
    x, hist, st = gpu.solve(method, s.b, s.x0, error=1e-11 * np.linalg.norm(s.b),
                            max_steps=5000)
    F = mesh.couplings_as_faces(s)
    gpu.set_system(F)
    y, _, _ = gpu.solve(method, s.b, s.x0, error=1e-11 * np.linalg.norm(s.b), max_steps=5000)
    np.testing.assert_allclose(x, y, rtol=1e-8, atol=1e-9)


def test_mixed_cg_with_couplings(gpu):
    s = _case(16, 4, True, shift=0.05)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 5000, 0)
    x, hist, it, rel, st = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 5000, 0)
    assert (it, st) == (rit, rst)
    np.testing.assert_array_equal(hist, rh)
    np.testing.assert_array_equal(x, rx)


@pytest.mark.parametrize("transpose", [0, 1])
def test_multi_rhs_bicg_with_couplings(gpu, transpose):
    s = _case(12, 5, False)
    gpu.set_system(s)
    A = O.ldu2csr(s)
    T = O.mirror_transpose(A)
    B = np.random.default_rng(6).uniform(-1, 1, (3, s.n_cells))
    X0 = np.zeros_like(B)
    X, its, nf, sts = gpu.solve_multi("BiCG", B, X0, error=0.0, max_steps=25,
                                      transpose=bool(transpose))
    for k in range(3):
        M, Mt = (T, A) if transpose else (A, T)
        ref = O.solve("BiCG", M, B[k], X0[k], 0.0, 25, At=Mt)
        np.testing.assert_array_equal(X[k], ref.x)


@pytest.mark.parametrize("name", ["cuFoamCG", "cuFoamBiCG", "cuFoamJMR", "cuFoamRichardson"])
def test_foam_wrappers_with_coupled_cyclic_interfaces(gpu, name):
    from cufoam_amd import foam
    sym = name in ("cuFoamCG", "cuFoamRichardson")
    s = _case(10, 7, sym)
    m = foam.lduMatrix.from_system(s)
    h = s.n_couples // 2
    itf = [foam.cyclicInterface(s.cpl_row[:h], s.cpl_col[:h]),
           foam.cyclicInterface(s.cpl_row[h:], s.cpl_col[h:])]
    bou = [-s.cpl_val[:h], -s.cpl_val[h:]]
    psi = np.random.default_rng(8).uniform(-1, 1, s.n_cells)
    psi_ref = psi.copy()
    solver = foam.New(name, "U", m, {"tolerance": 1e-9, "maxIter": 300,
                                     "interfaces": "coupled"}, gis=gpu,
                      interfaceBouCoeffs=bou, interfaceIntCoeffs=bou, interfaces=itf)
    perf = solver.solve(psi, s.b)
    ref_x, ref = O.foam_solve(name[len("cuFoam"):], s, psi_ref, s.b, 1e-9, 300)
    assert perf.n_iterations == ref.n_iterations
    assert perf.final_residual == ref.final_residual
    np.testing.assert_array_equal(psi, ref_x)
    # ignored interfaces (the reference's behaviour) solve the uncoupled matrix
    plain = foam.New(name, "U", m, {"tolerance": 1e-9, "maxIter": 300}, gis=gpu,
                     interfaceBouCoeffs=bou, interfaceIntCoeffs=bou, interfaces=itf)
    psi2 = psi_ref.copy()
    plain.solve(psi2, s.b)
    u = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, s.lower, s.upper, s.b,
                       s.x0)
    ref2_x, _ = O.foam_solve(name[len("cuFoam"):], u, psi_ref.copy(), s.b, 1e-9, 300)
    np.testing.assert_array_equal(psi2, ref2_x)


def test_interface_argument_checks_and_reset(gpu):
    from cufoam_amd.gis import GisError
    s = _case(4, 9, True)
    gpu.set_pattern(s.n_cells, s.lower_addr, s.upper_addr)
    plain_nnz = gpu.nnz
    with pytest.raises(GisError):
        gpu.set_interfaces([0, 1], [3, 0])            # (0,3) has no (3,0)
    with pytest.raises(GisError):
        gpu.set_interfaces([2], [2])                  # diagonal
    with pytest.raises(GisError):
        gpu.set_interfaces([0, s.n_cells], [s.n_cells, 0])
    gpu.set_interfaces(s.cpl_row, s.cpl_col)
    assert gpu.nnz == plain_nnz + s.n_couples
    with pytest.raises(GisError):
        gpu.set_coupled_values(s.diag, s.lower, s.upper, None)
    gpu.set_coupled_values(s.diag, s.lower, s.upper, s.cpl_val)
    gpu.set_interfaces([], [])
    assert gpu.nnz == plain_nnz
    gpu.set_interfaces(s.cpl_row, s.cpl_col)
    gpu.set_pattern(s.n_cells, s.lower_addr, s.upper_addr)   # clears the couplings
    assert gpu.nnz == plain_nnz
    # set_values after set_interfaces: coupling values are zero
    gpu.set_interfaces(s.cpl_row, s.cpl_col)
    gpu.set_values(s.diag, s.lower, s.upper)
    z = mesh.LduSystem(s.n_cells, s.lower_addr, s.upper_addr, s.diag, s.lower, s.upper, s.b,
                       s.x0, cpl_row=s.cpl_row, cpl_col=s.cpl_col,
                       cpl_val=np.zeros(s.n_couples))
    ref = O.solve("CG", O.ldu2csr(z), s.b, s.x0, 0.0, 10)
    x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=10)
    np.testing.assert_array_equal(hist, ref.hist)
