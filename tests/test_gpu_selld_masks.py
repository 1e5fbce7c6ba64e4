"""GPU parity of the SellD row masks in both widths.  A slice whose column-offset
table has at most 8 entries reads its rows' masks as one byte each, a larger
table reads 32-bit masks.  The system here is a 7-point block (tables of 6-7
entries) followed by a 27-point stencil block (tables of up to 27 entries), so
both widths occur, also within a pair of slices that one lane sums together.  SpMV
(FP64, FP32, transpose), CG and mixed-precision CG are compared bit for bit
with the oracle."""
import numpy as np
import pytest

from cufoam_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _stencil27(n, seed, shift):
    """SPD 27-point stencil on an n^3 grid: a face between every two cells whose
    i, j, k indices differ by at most 1, in upper-triangular order."""
    idx = np.arange(n ** 3).reshape(n, n, n)          # [k, j, i]
    owners, nbrs = [], []
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if (dk, dj, di) <= (0, 0, 0):
                    continue
                a = idx[max(0, -dk):n - max(0, dk), max(0, -dj):n - max(0, dj),
                        max(0, -di):n - max(0, di)]
                b = idx[max(0, dk):n + min(0, dk), max(0, dj):n + min(0, dj),
                        max(0, di):n + min(0, di)]
                lo, hi = np.minimum(a, b).ravel(), np.maximum(a, b).ravel()
                owners.append(lo)
                nbrs.append(hi)
    la = np.concatenate(owners)
    ua = np.concatenate(nbrs)
    order = np.lexsort((ua, la))
    la, ua = la[order].astype(np.int32), ua[order].astype(np.int32)
    rng = np.random.default_rng(seed)
    off = -(1.0 + 0.5 * rng.random(la.size))
    diag = mesh.row_abs_sum(n ** 3, la, ua, off, off) + shift
    b = rng.uniform(-1.0, 1.0, n ** 3)
    return mesh.LduSystem(n ** 3, la, ua, diag, off.copy(), off.copy(), b, np.zeros(n ** 3))


def _system(seed):
    """A 7-point 12^3 block (27 slices, tables of at most 7 entries) followed by
    a 27-point 16^3 block (tables of 18-27 entries): slices 26 and 27, one of
    each kind, are summed together by one lane."""
    a = mesh.laplacian(12, seed=seed + 1, shift=0.05)
    r = _stencil27(16, seed, 0.05)
    n0 = a.n_cells
    la = np.concatenate([a.lower_addr, r.lower_addr + n0]).astype(np.int32)
    ua = np.concatenate([a.upper_addr, r.upper_addr + n0]).astype(np.int32)
    return mesh.LduSystem(n0 + r.n_cells, la, ua, np.concatenate([a.diag, r.diag]),
                          np.concatenate([a.lower, r.lower]),
                          np.concatenate([a.upper, r.upper]),
                          np.concatenate([a.b, r.b]), np.zeros(n0 + r.n_cells))


def _table_sizes(A, n):
    """Offset-table size of each 64-row slice under the layout's rules (0: the
    slice keeps int32 columns): ascending rows of at most 32 entries, at most
    32 distinct offsets, padding at most a quarter."""
    rp, col = np.asarray(A.rowptr), np.asarray(A.col)
    sizes = []
    for s in range((n + 63) // 64):
        r0, r1 = s * 64, min(n, s * 64 + 64)
        offs, total, ok = set(), 0, True
        for r in range(r0, r1):
            c = col[rp[r]:rp[r + 1]]
            total += c.size
            ok &= c.size <= 32 and bool(np.all(np.diff(c) > 0))
            offs.update((c - r).tolist())
        u = len(offs)
        sizes.append(u if ok and u <= 32 and u * 64 * 4 <= 5 * total else 0)
    return np.array(sizes)


@pytest.fixture(scope="module")
def case():
    print("[case] 7-point 12^3 + 27-point 16^3 seed=31", flush=True)
    s = _system(31)
    A = O.ldu2csr(s)
    sizes = _table_sizes(A, s.n_cells)
    assert (sizes > 8).sum() >= 16 and ((sizes > 0) & (sizes <= 8)).sum() >= 16
    # a lane sums slices 2q and 2q + 1 together: one pair mixes the widths
    assert 0 < sizes[26] <= 8 < sizes[27]
    return s, A


def test_spmv_both_mask_widths(gpu, case):
    import torch
    s, A = case
    gpu.set_system(s)
    gpu.set_schedule(6, 6, -1)
    try:
        x = np.random.default_rng(32).standard_normal(s.n_cells)
        dx = torch.from_numpy(x).cuda()
        assert np.array_equal(gpu.spmv(dx).cpu().numpy(), O.spmv(A, x))
        assert np.array_equal(gpu.spmv(dx, transpose=True).cpu().numpy(),
                              O.spmv(O.transpose(A), x))
        x32 = dx.float()
        y32 = torch.empty_like(x32)
        gpu.spmv_async(x32, y32)
        gpu.synchronize()
        assert np.array_equal(y32.cpu().numpy(), O.spmv_f32(A, x.astype(np.float32)))
    finally:
        gpu.set_schedule(0, 0, -1)


def test_cg_and_mixed_both_mask_widths(gpu, case):
    s, A = case
    gpu.set_system(s)
    gpu.set_schedule(6, 6, -1)
    try:
        ref = O.solve("CG", A, s.b, s.x0, 0.0, 25)
        x, hist, st = gpu.solve("CG", s.b, s.x0, error=0.0, max_steps=25, hist_cap=32)
        np.testing.assert_array_equal(hist, ref.hist)
        np.testing.assert_array_equal(x, ref.x)
        rx, rh, rit, rst = O.mixed_cg(A, s.b, s.x0, 1e-10, 1e-2, 25, 0)
        x, hist, it, rel, st = gpu.solve_mixed(s.b, s.x0, 1e-10, 1e-2, 25, 0)
        assert (st, it) == (rst, rit)
        np.testing.assert_array_equal(hist, rh)
        np.testing.assert_array_equal(x, rx)
    finally:
        gpu.set_schedule(0, 0, -1)
