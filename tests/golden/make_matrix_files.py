"""Writes the Rutherford-Boeing and Matrix Market fixtures under tests/golden/
(data only: matrices built here from fixed seeds, stored in the column-compressed
layout the format defines, values printed with 17 significant digits).

    python tests/golden/make_matrix_files.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def grid_matrix(nx, ny, seed, symmetric):
    """5-point operator on an nx x ny grid with random weights, diagonally dominant."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    A = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            c = i + nx * j
            for di, dj in ((1, 0), (0, 1)):
                if i + di < nx and j + dj < ny:
                    d = c + di + nx * dj
                    w = -(0.5 + rng.random())
                    A[c, d] = w
                    A[d, c] = w if symmetric else w * (0.5 + rng.random())
    A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=1) + 0.25
    return A


def write_rb(path, A, mtype, title):
    """Column-compressed storage of A (lower triangle for s/z types), 1-based."""
    nrow, ncol = A.shape
    ptr, idx, val = [1], [], []
    for j in range(ncol):
        rows = np.nonzero(A[:, j])[0]
        if mtype[1] in "sz":
            rows = rows[rows >= j] if mtype[1] == "s" else rows[rows > j]
        for r in rows:
            idx.append(r + 1)
            val.append(A[r, j])
        ptr.append(len(idx) + 1)
    nnz = len(idx)

    def block(items, per, fmt):
        return ["".join(fmt % v for v in items[k:k + per]) for k in range(0, len(items), per)]

    pl = block(ptr, 10, "%8d")
    il = block(idx, 10, "%8d")
    vl = block(val, 3, "%25.17E")
    lines = [f"{title:<72s}{'GISFIX':>8s}",
             f"{len(pl) + len(il) + len(vl):14d}{len(pl):14d}{len(il):14d}{len(vl):14d}",
             f"{mtype:<3s}{'':11s}{nrow:14d}{ncol:14d}{nnz:14d}{0:14d}",
             f"{'(10I8)':<16s}{'(10I8)':<16s}{'(3E25.17)':<20s}"]
    with open(path, "w") as f:
        f.write("\n".join(lines + pl + il + vl) + "\n")


def main():
    sym = grid_matrix(6, 6, 1, True)
    uns = grid_matrix(5, 5, 2, False)
    rect = np.random.default_rng(3).standard_normal((4, 6)) * (np.random.default_rng(4).random((4, 6)) < 0.6)
    skew = np.zeros((5, 5))
    r = np.random.default_rng(5)
    for i in range(5):
        for j in range(i + 1, 5):
            if r.random() < 0.6:
                skew[j, i] = r.standard_normal()
                skew[i, j] = -skew[j, i]
    write_rb(os.path.join(HERE, "grid6_rsa.rb"), sym, "rsa", "6x6 grid, symmetric weights")
    write_rb(os.path.join(HERE, "grid5_rua.rb"), uns, "rua", "5x5 grid, unsymmetric weights")
    write_rb(os.path.join(HERE, "rect4x6_rra.rb"), rect, "rra", "4 x 6 rectangular")
    write_rb(os.path.join(HERE, "skew5_rza.rb"), skew, "rza", "5 x 5 skew-symmetric")
    write_rb(os.path.join(HERE, "grid6_psa.rb"), sym, "psa", "pattern only (unsupported)")
    write_rb(os.path.join(HERE, "grid6_rse.rb"), sym, "rse", "elemental (unsupported)")
    np.savez(os.path.join(HERE, "matrix_files_dense.npz"), grid6_rsa=sym, grid5_rua=uns,
             rect4x6_rra=rect, skew5_rza=skew)
    v = np.random.default_rng(6).standard_normal(7)
    with open(os.path.join(HERE, "vec7.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix array real general\n7 1\n")
        f.write("".join("%.17E\n" % x for x in v))
    np.save(os.path.join(HERE, "vec7.npy"), v)


if __name__ == "__main__":
    main()
