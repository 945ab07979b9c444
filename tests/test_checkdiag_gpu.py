"""nrg_check_diag_batched: the reference's check_diag (c++/numerics.hpp:340-404: eigenvalues finite,
eigenvector rows orthonormal within 1e-12 / 1e-10) for a whole step on the device.

The reported maxima must equal max |V V^T - I| formed by numpy from the same vectors; planted defects
(a row scaled by 1 + 1e-9, a duplicated row, a NaN eigenvalue) must be reported with their codes for
their block only."""
import numpy as np
import pytest

from nrgljubljana_amd import NRG_SYEV_FP64, workloads

pytestmark = pytest.mark.gpu


def pack(blocks):
    sizes = np.array([b.shape[0] for b in blocks], np.int64)
    off_H = np.zeros(len(sizes), np.int64)
    off_w = np.zeros(len(sizes), np.int64)
    off_H[1:] = np.cumsum(sizes[:-1] ** 2)
    off_w[1:] = np.cumsum(sizes[:-1])
    return sizes, off_H, np.concatenate([b.ravel() for b in blocks]), off_w


def solve(ctx, blocks):
    import torch
    sizes, off_H, arena, off_w = pack(blocks)
    H = torch.from_numpy(arena).cuda()
    w = torch.zeros(int(sizes.sum()), dtype=torch.float64, device="cuda")
    ctx.syev_batched(H, sizes, off_H, w, off_w, NRG_SYEV_FP64)
    return sizes, off_H, off_w, H, w


def gram_dev(V, m):
    G = V[:m] @ V[:m].T
    return float(np.max(np.abs(np.diag(G) - 1.0))), float(np.max(np.abs(G - np.diag(np.diag(G)))) if m > 1 else 0.0)


def test_valid_eigenvectors_pass_with_exact_maxima(ctx):
    rng = np.random.default_rng(31)
    orders = [1, 2, 17, 64, 65, 200, 700, 1300]
    blocks = [workloads.eig_block(rng, n)[0] for n in orders]
    sizes, off_H, off_w, H, w = solve(ctx, blocks)
    rows = sizes.copy()
    rows[4] = 30   # a dsyevr-like partial set of eigenvectors
    info, dn, do = ctx.check_diag_batched(H, sizes, rows, off_H, w, off_w)
    assert not info.any(), info
    a = H.cpu().numpy()
    for b, n in enumerate(sizes):
        V = a[off_H[b]:off_H[b] + n * n].reshape(n, n)
        en, eo = gram_dev(V, int(rows[b]))
        print(f"n={n} rows={rows[b]}: norm {dn[b]:.2e} (numpy {en:.2e}), orth {do[b]:.2e} (numpy {eo:.2e})", flush=True)
        # the device Gram matrix is an FP64 product of the same vectors: equal to a few ulps of n eps
        assert abs(dn[b] - en) <= 4 * n * 2.2e-16 and abs(do[b] - eo) <= 4 * n * 2.2e-16
        assert dn[b] <= 2e-12 and do[b] <= 1e-10


def test_planted_defects_are_reported_per_block(ctx):
    rng = np.random.default_rng(32)
    orders = [40, 300, 90, 500]
    blocks = [workloads.eig_block(rng, n)[0] for n in orders]
    sizes, off_H, off_w, H, w = solve(ctx, blocks)
    Vs = H.view(-1)
    n0, n1, n3 = int(sizes[0]), int(sizes[1]), int(sizes[3])
    Vs[off_H[0] + 5 * n0: off_H[0] + 6 * n0] *= 1 + 1e-9                 # block 0: |<v5, v5> - 1| = 2e-9
    Vs[off_H[1] + 7 * n1: off_H[1] + 8 * n1] = Vs[off_H[1] + 3 * n1: off_H[1] + 4 * n1].clone()   # block 1: v7 = v3
    w[int(off_w[3]) + 11] = float("nan")                                   # block 3: non-finite eigenvalue
    info, dn, do = ctx.check_diag_batched(H, sizes, sizes, off_H, w, off_w)
    print("info", info, "norm", dn, "orth", do, flush=True)
    assert list(info) == [1, 2, 0, 3]
    assert dn[0] == pytest.approx(2e-9, rel=1e-3)
    assert do[1] == pytest.approx(1.0, abs=1e-12)
    assert dn[2] <= 2e-12 and do[2] <= 1e-10


def test_nan_eigenvector_counts_as_infinite(ctx):
    rng = np.random.default_rng(33)
    blocks = [workloads.eig_block(rng, 120)[0]]
    sizes, off_H, off_w, H, w = solve(ctx, blocks)
    H[17] = float("nan")
    info, dn, do = ctx.check_diag_batched(H, sizes, sizes, off_H, w, off_w)
    assert info[0] == 1 and np.isinf(dn[0]) and np.isinf(do[0])
