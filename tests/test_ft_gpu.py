"""GPU parity of nrg_ft_bin against the oracle (Algo_FT::calc, c++/algo_FT.hpp:28-44, binned by
ChainBinning::add, c++/spectrum.hpp:19-28).

Tolerance: the device exp and the atomics' order of addition. Every bin must agree to 1e-13 of the
summed absolute weight of its array."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, grid, np_ptr, random_pair, random_step
from tests.test_fdm_gpu import device_structs

pytestmark = pytest.mark.gpu
T = 0.5


def gpu_run(ctx, pairs, g, sign=-1, Z=1.0, out=None):
    import torch
    structs, keep = device_structs(pairs)
    if out is None:
        out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(2)]
    ctx.ft_bin(structs, g, T, Z, sign, *out)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def check(ctx, pairs, g, sign=-1, Z=1.0, tol=1e-13):
    print(f"ft parity: {len(pairs)} pairs, nbins {g.nbins}, sign {sign}", flush=True)
    got = gpu_run(ctx, pairs, g, sign, Z)
    ref = oracle.ft_bin([p.to_struct(np_ptr) for p in pairs], g, T, Z, sign, g.nbins)
    ab = [Pair(p.Ei, p.Ej, np.abs(p.op1[:, :p.di]), np.abs(p.op2[:, :p.di]), p.Ki, p.Kj, factor=abs(p.factor))
          for p in pairs]
    scale = [max(np.abs(a).sum(), 1e-300) for a in oracle.ft_bin([p.to_struct(np_ptr) for p in ab], g, T, Z, -1,
                                                                 g.nbins)]
    for name, a, b, s in zip(("pos", "neg"), got, ref, scale):
        err = float(np.max(np.abs(a - b)))
        assert err <= tol * s, f"{name}: max bin error {err:.3e} vs scale {s:.3e}"


def test_hand_value_and_sum_rule(ctx):
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, factor=0.5)
    pos, neg = gpu_run(ctx, [p], g, Z=4.0)
    assert pos[20] == pytest.approx(6.0 * (math.exp(-0.3 / T) + math.exp(-0.2 / T)) / 4.0 * 0.5, rel=1e-15)
    rng = np.random.default_rng(3)
    m = 300
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    Ei, Ej = np.sort(rng.uniform(0, 2, m)), np.sort(rng.uniform(0, 2, m))
    Z = np.exp(-Ei / T).sum() + np.exp(-Ej / T).sum()
    pos, neg = gpu_run(ctx, [Pair(Ei, Ej, d, d, m, m)], grid(-8.0, 100.0, log10emax=1.0), Z=Z)
    assert pos.sum() + neg.sum() == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("seed", [1, 2])
def test_random_steps_match_oracle(ctx, seed):
    pairs = random_step(seed, npairs=24, dmax=300)
    check(ctx, pairs, grid(-6.0, 50.0, log10emax=1.0), Z=7.5)
    check(ctx, pairs, grid(-6.0, 50.0, log10emax=1.0), sign=1, Z=7.5)


def test_large_pair_and_bad_arguments(ctx):
    import torch
    rng = np.random.default_rng(11)
    check(ctx, [random_pair(rng, 1100, 900, 800, 700, pad=5)], grid(-8.0, 1000.0, log10emax=1.0), Z=40.0)
    g = grid()
    out = [torch.zeros(g.nbins, dtype=torch.float64, device="cuda") for _ in range(2)]
    structs, keep = device_structs([random_pair(rng, 5, 4, 2, 3)])
    for kw in (dict(Z=0.0), dict(sign=0)):
        args = dict(Z=1.0, sign=-1)
        args.update(kw)
        with pytest.raises(NrgError):
            ctx.ft_bin(structs, g, T, args["Z"], args["sign"], *out)
    ctx.ft_bin(structs, g, T, 1.0, -1, *out)
    torch.cuda.synchronize()
