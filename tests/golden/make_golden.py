"""Write the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture is data only: the LDU system (addressing, coefficients, b, x0),
the expected CSR (row offsets, columns, values) and, per method, the residual
history and solution after a fixed number of iterations, all float64/int32.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cufoam_amd import mesh  # noqa: E402
from oracle import oracle as O  # noqa: E402

METHODS = ["CG", "BiCG", "Jacobi", "MR", "MRJacobi", "JMR", "MRS", "Richardson"]
STEPS = 20


def systems():
    yield "hex4_nonsym", mesh.laplacian(4, seed=1, shift=0.1, symmetric=False)
    yield "hex8_sym", mesh.laplacian(8, seed=0, shift=1e-3, symmetric=True)
    yield "random300_nonsym", mesh.random_ldu(300, 900, seed=3)
    yield "random300_sym", mesh.random_ldu(300, 900, seed=4, symmetric=True)


def fixture(name, s):
    A = O.ldu2csr(s)
    out = dict(n_cells=np.int32(s.n_cells), lower_addr=s.lower_addr, upper_addr=s.upper_addr,
               diag=s.diag, lower=s.lower, upper=s.upper, b=s.b,
               x0=np.random.default_rng(11).uniform(-1, 1, s.n_cells),
               rowptr=A.rowptr, col=A.col, val=A.val)
    for m in METHODS:
        r = O.solve(m, A, s.b, out["x0"], 0.0, STEPS)
        out[f"{m}_hist"] = r.hist
        out[f"{m}_x"] = r.x
    if name == "hex8_sym":
        s2 = mesh.laplacian(8, seed=0, shift=1e-4, symmetric=True)
        A2 = O.ldu2csr(s2)
        for mode in (0, 1):
            x, h, it, st = O.mixed_cg(A2, s2.b, s2.x0, 1e-10, 1e-2, 5000, mode)
            out[f"mixed{mode}_diag"] = s2.diag
            out[f"mixed{mode}_x"] = x
            out[f"mixed{mode}_hist"] = h
            out[f"mixed{mode}_iters"] = np.int32(it)
    return out


def main():
    for name, s in systems():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fixture(name, s))
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
