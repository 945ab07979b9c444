"""GPU parity of nrg_fdm_bin against the oracle (Algo_FDM::calc + ChainBinning, c++/algo_FDM.hpp:30-60,
84-114,136-142, c++/bins.hpp:103-132).

Tolerance: the kernel evaluates each weight in the reference's order, so the only differences are the
device exp/log10 (a few ulp) and the order of the atomic additions into a bin. Every bin must agree
to 1e-13 of the summed absolute weight of its array (a weight moved to a neighbour bin by a one-ulp
change of log10 moves by at most that much, since the split is continuous in x)."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, grid, np_ptr, random_pair, random_step

pytestmark = pytest.mark.gpu

T = 0.1
WNF = 0.6


def device_structs(pairs):
    import torch
    keep = {}

    def ptr(a):
        if id(a) not in keep:   # the array is held too, so its id cannot be reused meanwhile
            keep[id(a)] = (a, torch.from_numpy(a).cuda())
        return keep[id(a)][1].data_ptr()

    return [p.to_struct(ptr) for p in pairs], keep


def gpu_run(ctx, pairs, g, sign=-1, T=T, wnf=WNF, out=None):
    import torch
    structs, keep = device_structs(pairs)
    if out is None:
        out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    ctx.fdm_bin(structs, g, T, wnf, sign, *out)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def cpu_run(pairs, g, sign=-1, T=T, wnf=WNF):
    return oracle.fdm_bin([p.to_struct(np_ptr) for p in pairs], g, T, wnf, sign, g.nbins)


def abs_scale(pairs, g, sign=-1):
    """Summed absolute weight per output array (oracle with |op1|, |op2|, |rho|)."""
    ab = [Pair(p.Ei, p.Ej, np.abs(p.op1[:, :p.di]), np.abs(p.op2[:, :p.di]), p.Ki, p.Kj,
               None if p.rho_i is None else np.abs(p.rho_i), None if p.rho_j is None else np.abs(p.rho_j),
               abs(p.factor)) for p in pairs]
    return [max(np.abs(a).sum(), 1e-300) for a in cpu_run(ab, g, sign)]


def check(ctx, pairs, g, sign=-1, tol=1e-13):
    print(f"fdm parity: {len(pairs)} pairs, nbins {g.nbins}, sign {sign}", flush=True)
    got = gpu_run(ctx, pairs, g, sign)
    ref = cpu_run(pairs, g, sign)
    scale = abs_scale(pairs, g, sign)
    for name, a, b, s in zip(("gt_pos", "gt_neg", "ls_pos", "ls_neg"), got, ref, scale):
        err = float(np.max(np.abs(a - b))) if a.size else 0.0
        assert err <= tol * s, f"{name}: max bin error {err:.3e} vs scale {s:.3e}"
    return got, ref


def test_single_peak_matches_hand_value(ctx):
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.0], [0.1], np.array([[2.0]]), np.array([[3.0]]), 0, 0, factor=0.5)
    gtp, gtn, lsp, lsn = gpu_run(ctx, [p], g)
    assert gtp[20] == pytest.approx(0.5 * 6.0 * WNF, rel=1e-15)
    assert lsp[20] == pytest.approx(0.5 * 6.0 * math.exp(-1.0) * WNF, rel=1e-15)
    assert np.count_nonzero(gtp) == 1 and np.count_nonzero(lsp) == 1 and not gtn.any() and not lsn.any()


def test_density_matrix_products_hand_values(ctx):
    g = grid(-3.0, 10.0, log10emax=1.0)
    rho = np.array([[7.0, 11.0], [13.0, 17.0]])
    p = Pair([0.0, 0.0], [0.1], np.array([[1.0, 2.0]]), np.array([[3.0, 5.0]]), 2, 0, rho_i=rho, factor=2.0)
    gtp, _, _, _ = gpu_run(ctx, [p], g)
    assert gtp[20] == 2.0 * (86 + 2 * 118)
    q = Pair([0.0], [0.1, 0.1], np.array([[1.0], [2.0]]), np.array([[3.0], [5.0]]), 0, 2, rho_j=rho)
    _, _, lsp, _ = gpu_run(ctx, [q], g, sign=1)
    assert lsp[20] == -(76 + 2 * 124)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_steps_match_oracle(ctx, seed):
    check(ctx, random_step(seed, npairs=24, dmax=300), grid(-6.0, 50.0, log10emax=1.0))


def test_fermion_and_boson(ctx):
    pairs = random_step(7, npairs=6, dmax=100)
    g = grid(-5.0, 20.0, log10emax=1.0)
    check(ctx, pairs, g, sign=-1)
    check(ctx, pairs, g, sign=1)


def test_large_pair_and_many_bins(ctx):
    rng = np.random.default_rng(11)
    pairs = [random_pair(rng, 1500, 1300, 700, 500, pad=5), random_pair(rng, 900, 1100, 900, 0)]
    check(ctx, pairs, grid(-8.0, 1000.0, log10emax=1.0))


def test_degenerate_energies_and_clamps(ctx):
    rng = np.random.default_rng(4)
    p = random_pair(rng, 90, 70, 30, 20)
    p.Ej[:] = np.round(p.Ej, 1)   # many exact zero-energy differences after the shift
    p.Ei[:] = np.round(p.Ei, 1)
    check(ctx, [p], grid(-2.0, 10.0, log10emax=0.0))


def test_discard_immediately(ctx):
    pairs = random_step(5, npairs=8, dmax=120)
    check(ctx, pairs, grid(-6.0, 30.0, log10emax=1.0, discard=1e-2))


def test_accumulates_and_empty_call(ctx):
    import torch
    pairs = random_step(9, npairs=4, dmax=80)
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    once = [a.copy() for a in gpu_run(ctx, pairs, g, out=out)]
    twice = gpu_run(ctx, pairs, g, out=out)
    # atomics add in any order, so a bin with cancellation differs by rounding of its terms
    for a, b, s in zip(once, twice, abs_scale(pairs, g)):
        np.testing.assert_allclose(b, 2 * a, rtol=0, atol=1e-13 * s)
    ctx.fdm_bin([], g, T, WNF, -1, *out)
    torch.cuda.synchronize()
    for a, b in zip(twice, out):
        np.testing.assert_array_equal(a, b.cpu().numpy())


def test_bad_arguments(ctx):
    import torch
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(4)]
    rng = np.random.default_rng(0)
    good = random_pair(rng, 5, 4, 2, 1)
    structs, keep = device_structs([good])
    s = structs[0]

    def expect_fail(pairs, grid_=g, T_=T, sign=-1):
        with pytest.raises(NrgError):
            ctx.fdm_bin(pairs, grid_, T_, WNF, sign, *out)

    bad = type(s).from_buffer_copy(s); bad.K_i = 6
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.rho_i = None
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.n_rho_j = 0
    expect_fail([bad])
    bad = type(s).from_buffer_copy(s); bad.ld_op = 3
    expect_fail([bad])
    expect_fail([s], T_=0.0)
    expect_fail([s], sign=0)
    expect_fail([s], grid_=grid(nbins=1))
    # the context stays usable
    ctx.fdm_bin([s], g, T, WNF, -1, *out)
    torch.cuda.synchronize()
