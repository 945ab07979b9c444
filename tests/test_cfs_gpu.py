"""GPU parity of nrg_cfs_bin against the oracle (Algo_CFS::calc, c++/algo_CFS.hpp:29-56, 79-106,
binned by ChainBinning::add, c++/spectrum.hpp:19-28).

Tolerance: the density-matrix products are formed on the FP64 matrix cores in another order than
the oracle's dgemm, the Boltzmann factors use the device exp, and the bins are filled by atomics in
any order. Every bin must agree to 1e-13 of the summed absolute weight of its array."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, grid, np_ptr, random_pair, random_step
from tests.test_fdm_gpu import device_structs

pytestmark = pytest.mark.gpu
T = 0.5


def gpu_run(ctx, pairs, g, last, sign=-1, Z=1.0, out=None):
    import torch
    structs, keep = device_structs(pairs)
    if out is None:
        out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    ctx.cfs_bin(structs, g, T, Z, last, sign, *out)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def cpu_run(pairs, g, last, sign=-1, Z=1.0):
    return oracle.cfs_bin([p.to_struct(np_ptr) for p in pairs], g, T, Z, last, sign, g.nbins)


def abs_scale(pairs, g, last, Z):
    ab = [Pair(p.Ei, p.Ej, np.abs(p.op1[:, :p.di]), np.abs(p.op2[:, :p.di]), p.Ki, p.Kj,
               None if p.rho_i is None else np.abs(p.rho_i), None if p.rho_j is None else np.abs(p.rho_j),
               abs(p.factor)) for p in pairs]
    return [max(np.abs(a).sum(), 1e-300) for a in cpu_run(ab, g, last, -1, Z)]


def cfs_step(seed, npairs, dmax):
    """random_step with every density matrix inside its subspace (n_rho <= dim, prod_fit_left's bound)."""
    pairs = random_step(seed, npairs=npairs, dmax=dmax)
    for p in pairs:
        if p.rho_i is not None and p.rho_i.shape[0] > p.di:
            p.rho_i = np.ascontiguousarray(p.rho_i[:p.di, :p.di])
        if p.rho_j is not None and p.rho_j.shape[0] > p.dj:
            p.rho_j = np.ascontiguousarray(p.rho_j[:p.dj, :p.dj])
    return pairs


def check(ctx, pairs, g, last, sign=-1, Z=1.0, tol=1e-13):
    print(f"cfs parity: {len(pairs)} pairs, nbins {g.nbins}, last {last}, sign {sign}", flush=True)
    got = gpu_run(ctx, pairs, g, last, sign, Z)
    ref = cpu_run(pairs, g, last, sign, Z)
    for name, a, b, s in zip(("gt_pos", "gt_neg", "ls_pos", "ls_neg"), got, ref, abs_scale(pairs, g, last, Z)):
        err = float(np.max(np.abs(a - b))) if a.size else 0.0
        assert err <= tol * s, f"{name}: max bin error {err:.3e} vs scale {s:.3e}"


def test_hand_values(ctx):
    # the 2 x 2 pair of the oracle's pins: 13.5 in ls bin 30 (Eq. 16), 6 in gt_neg at 0.25 (Eq. 15)
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.125, 0.5], [0.25, 1.125], np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([[5.0, 6.0], [7.0, 8.0]]),
             1, 1, np.array([[0.5, 0.25], [0.125, 0.0625]]), np.array([[0.3, 0.2], [0.1, 0.4]]))
    gp, gn, lp, ln = gpu_run(ctx, [p], g, False)
    assert lp[30] == 13.5 and np.count_nonzero(lp) == 1 and not ln.any() and not gp.any()
    assert gn.sum() == pytest.approx(6.0, rel=1e-15)
    gp, gn, lp, ln = gpu_run(ctx, [p], g, True, Z=2.0)
    assert gp.sum() == pytest.approx(5.0 * math.exp(-0.125 / T) / 2.0, rel=1e-15)
    assert lp.sum() == pytest.approx(5.0 * math.exp(-0.25 / T) / 2.0, rel=1e-15)


def test_sum_rules(ctx):
    rng = np.random.default_rng(3)
    m = 300
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    a = rng.standard_normal((m, m))
    ri = a @ a.T
    ri /= np.trace(ri)
    E = (np.sort(rng.uniform(0, 1, m)), np.sort(rng.uniform(0, 1, m)))
    g = grid(-8.0, 100.0, log10emax=1.0)
    out = gpu_run(ctx, [Pair(*E, d, d, m, 0, ri, None)], g, False)
    assert out[2].sum() + out[3].sum() == pytest.approx(1.0, abs=1e-12)
    out = gpu_run(ctx, [Pair(*E, d, d, 0, m, None, ri)], g, False)
    assert out[0].sum() + out[1].sum() == pytest.approx(1.0, abs=1e-12)
    Z = np.exp(-E[0] / T).sum() + np.exp(-E[1] / T).sum()
    out = gpu_run(ctx, [Pair(*E, d, d, m, m)], g, True, Z=Z)
    assert sum(o.sum() for o in out) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_steps_match_oracle(ctx, seed):
    pairs = cfs_step(seed, 24, 300)
    g = grid(-6.0, 50.0, log10emax=1.0)
    check(ctx, pairs, g, False)
    check(ctx, pairs, g, True, Z=7.5)


def test_boson_sign(ctx):
    pairs = cfs_step(7, 8, 150)
    g = grid(-5.0, 20.0, log10emax=1.0)
    check(ctx, pairs, g, False, sign=1)
    check(ctx, pairs, g, True, sign=1, Z=3.0)


def test_large_pairs_many_bins_and_discard(ctx):
    rng = np.random.default_rng(11)
    pairs = [random_pair(rng, 1100, 900, 500, 300, pad=5), random_pair(rng, 500, 600, 440, 100),
             random_pair(rng, 700, 650, 0, 650), random_pair(rng, 640, 700, 640, 0)]
    for last in (False, True):
        check(ctx, pairs, grid(-8.0, 1000.0, log10emax=1.0), last, Z=50.0)
        check(ctx, pairs, grid(-8.0, 100.0, log10emax=1.0, discard=1e-2), last, Z=50.0)


def test_accumulates_and_empty_call(ctx):
    import torch
    pairs = cfs_step(9, 4, 80)
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    once = [a.copy() for a in gpu_run(ctx, pairs, g, False, out=out)]
    twice = gpu_run(ctx, pairs, g, False, out=out)
    for a, b, s in zip(once, twice, abs_scale(pairs, g, False, 1.0)):
        np.testing.assert_allclose(b, 2 * a, rtol=0, atol=1e-13 * s)
    ctx.cfs_bin([], g, T, 1.0, False, -1, *out)
    torch.cuda.synchronize()
    for a, b in zip(twice, out):
        np.testing.assert_array_equal(a, b.cpu().numpy())


def test_bad_arguments(ctx):
    import torch
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    rng = np.random.default_rng(0)
    good = random_pair(rng, 5, 4, 2, 3)
    structs, keep = device_structs([good])
    s = structs[0]

    def expect_fail(pairs, last=False, Z=1.0, sign=-1, grid_=g):
        with pytest.raises(NrgError):
            ctx.cfs_bin(pairs, grid_, T, Z, last, sign, *out)

    bad = type(s).from_buffer_copy(s); bad.rho_j = None
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.n_rho_i = 1
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.n_rho_i = 6      # beyond dim_i (prod_fit_left's bound)
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.K_j = 5
    expect_fail([bad])
    expect_fail([s], last=True, Z=0.0)
    expect_fail([s], sign=2)
    expect_fail([s], grid_=grid(nbins=1))
    bad = type(s).from_buffer_copy(s); bad.rho_i = None; bad.rho_j = None
    ctx.cfs_bin([bad], g, T, 1.0, True, -1, *out)      # the last step reads no density matrix
    ctx.cfs_bin([s], g, T, 1.0, False, -1, *out)
    torch.cuda.synchronize()
