"""GPU: matrices read from cuFoam's file formats run through the LDU device path
and give the oracle's results on the file's own CSR, bit for bit."""
import os

import numpy as np
import pytest

from cufoam_amd import matrix_io as mio
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,method", [("grid6_rsa", "CG"), ("grid6_rsa", "Richardson"),
                                         ("grid5_rua", "BiCG"), ("grid5_rua", "MRJacobi"),
                                         ("grid5_rua", "JMR")])
def test_file_matrix_solves_bit_exact(gpu, name, method, tmp_path):
    A = mio.read(os.path.join(GOLD, name + ".rb"))
    p = str(tmp_path / "m.csr")            # and through the binary format
    mio.save_csr(p, A)
    A = mio.load_csr(p)
    b = np.random.default_rng(3).uniform(-1, 1, A.num_rows)
    L = mio.csr_to_ldu(A, b=b)
    print(f"[case] {name} {method} N={A.num_rows} nnz={A.num_entries}", flush=True)
    gpu.set_system(L)
    rp, col, val = gpu.get_csr()
    np.testing.assert_array_equal(rp, A.row_offsets)
    np.testing.assert_array_equal(col, A.column_indices)
    np.testing.assert_array_equal(val, A.values)
    Ao = O.Csr(A.num_rows, A.row_offsets, A.column_indices, A.values)
    err = 1e-10 * np.linalg.norm(b)
    ref = O.solve(method, Ao, b, np.zeros(A.num_rows), err, 500)
    x, hist, st = gpu.solve(method, b, np.zeros(A.num_rows), error=err, max_steps=500,
                            hist_cap=ref.hist_len + 4)
    assert st == ref.status
    np.testing.assert_array_equal(hist, ref.hist)
    np.testing.assert_array_equal(x, ref.x)
