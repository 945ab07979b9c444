"""GPU parity of nrg_dmnrg_bin against the oracle (Algo_DMNRG::calc, c++/algo_DMNRG.hpp:30-52, binned by
ChainBinning::add, c++/spectrum.hpp:19-28).

Tolerance: the two density-matrix sums of every element are formed on the FP64 matrix cores in
another order than the oracle's loops, and the bins are filled by atomics in any order. Every bin
must agree to 1e-13 of the summed absolute weight of its array."""
import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, grid, np_ptr, random_pair, random_step
from tests.test_fdm_gpu import device_structs

pytestmark = pytest.mark.gpu

INF = float("inf")


def gpu_run(ctx, pairs, g, emin=0.0, emax=INF, sign=-1, out=None):
    import torch
    structs, keep = device_structs(pairs)
    if out is None:
        out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(2)]
    ctx.dmnrg_bin(structs, g, emin, emax, sign, *out)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def cpu_run(pairs, g, emin=0.0, emax=INF, sign=-1):
    return oracle.dmnrg_bin([p.to_struct(np_ptr) for p in pairs], g, emin, emax, sign, g.nbins)


def abs_scale(pairs, g, emin, emax, sign):
    ab = [Pair(p.Ei, p.Ej, np.abs(p.op1[:, :p.di]), np.abs(p.op2[:, :p.di]), p.Ki, p.Kj,
               None if p.rho_i is None else np.abs(p.rho_i), None if p.rho_j is None else np.abs(p.rho_j),
               abs(p.factor)) for p in pairs]
    return [max(np.abs(a).sum(), 1e-300) for a in cpu_run(ab, g, emin, emax, -1)]


def dm_step(seed, npairs, dmax):
    """random_step with density matrices on every pair that has kept states on both sides."""
    pairs = random_step(seed, npairs=npairs, dmax=dmax)
    rng = np.random.default_rng(seed + 100)
    for p in pairs:
        if p.Ki and p.Kj:
            if p.rho_i is None or p.rho_i.shape[0] < p.Ki:
                p.rho_i = rng.standard_normal((p.Ki, p.Ki)) / p.Ki
            if p.rho_j is None or p.rho_j.shape[0] < p.Kj:
                p.rho_j = rng.standard_normal((p.Kj, p.Kj)) / p.Kj
    return pairs


def check(ctx, pairs, g, emin=0.0, emax=INF, sign=-1, tol=1e-13):
    print(f"dmnrg parity: {len(pairs)} pairs, nbins {g.nbins}, window [{emin}, {emax}], sign {sign}", flush=True)
    got = gpu_run(ctx, pairs, g, emin, emax, sign)
    ref = cpu_run(pairs, g, emin, emax, sign)
    for name, a, b, s in zip(("pos", "neg"), got, ref, abs_scale(pairs, g, emin, emax, sign)):
        err = float(np.max(np.abs(a - b))) if a.size else 0.0
        assert err <= tol * s, f"{name}: max bin error {err:.3e} vs scale {s:.3e}"


def test_single_element_hand_value(ctx):
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]),
             factor=0.5)
    pos, neg = gpu_run(ctx, [p], g)
    assert pos[20] == 0.5 * (3.0 * 0.25 * 2.0 + 2.0 * 0.5 * 3.0)
    assert np.count_nonzero(pos) == 1 and not neg.any()


def test_fermion_sum_rule(ctx):
    rng = np.random.default_rng(3)
    m = 300
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    a = rng.standard_normal((m, m))
    b = rng.standard_normal((m, m))
    ri, rj = a @ a.T, b @ b.T
    t = np.trace(ri) + np.trace(rj)
    p = Pair(rng.uniform(0, 1, m), rng.uniform(0, 1, m), d, d, m, m, ri / t, rj / t)
    pos, neg = gpu_run(ctx, [p], grid(-8.0, 100.0, log10emax=1.0))
    assert pos.sum() + neg.sum() == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_steps_match_oracle(ctx, seed):
    check(ctx, dm_step(seed, 24, 300), grid(-6.0, 50.0, log10emax=1.0))


def test_energy_window_and_boson_sign(ctx):
    pairs = dm_step(7, 8, 150)
    g = grid(-5.0, 20.0, log10emax=1.0)
    check(ctx, pairs, g, emin=0.05, emax=0.6, sign=-1)
    check(ctx, pairs, g, emin=0.05, emax=0.6, sign=1)


def test_large_pair_many_bins_and_discard(ctx):
    rng = np.random.default_rng(11)
    pairs = [random_pair(rng, 1100, 900, 800, 700, pad=5), random_pair(rng, 500, 600, 440, 600)]
    check(ctx, pairs, grid(-8.0, 1000.0, log10emax=1.0))
    check(ctx, pairs, grid(-8.0, 100.0, log10emax=1.0, discard=1e-2))


def test_accumulates_and_empty_call(ctx):
    import torch
    pairs = dm_step(9, 4, 80)
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(2)]
    once = [a.copy() for a in gpu_run(ctx, pairs, g, out=out)]
    twice = gpu_run(ctx, pairs, g, out=out)
    for a, b, s in zip(once, twice, abs_scale(pairs, g, 0.0, INF, -1)):
        np.testing.assert_allclose(b, 2 * a, rtol=0, atol=1e-13 * s)
    ctx.dmnrg_bin([], g, 0.0, INF, -1, *out)
    torch.cuda.synchronize()
    for a, b in zip(twice, out):
        np.testing.assert_array_equal(a, b.cpu().numpy())


def test_bad_arguments(ctx):
    import torch
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(2)]
    rng = np.random.default_rng(0)
    good = random_pair(rng, 5, 4, 2, 3)
    structs, keep = device_structs([good])
    s = structs[0]

    def expect_fail(pairs, emin=0.0, emax=INF, sign=-1, grid_=g):
        with pytest.raises(NrgError):
            ctx.dmnrg_bin(pairs, grid_, emin, emax, sign, *out)

    bad = type(s).from_buffer_copy(s); bad.rho_j = None
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.n_rho_i = 1
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.K_j = 5
    expect_fail([bad])
    expect_fail([s], emin=1.0, emax=0.5)
    expect_fail([s], sign=2)
    expect_fail([s], grid_=grid(nbins=1))
    ctx.dmnrg_bin([s], g, 0.0, INF, -1, *out)
    torch.cuda.synchronize()
