"""GPU parity of nrg_temp_sum against the oracle (Algo_GT<n> and Algo_CHIT, c++/algo_FT.hpp:95-170).

Tolerance: the kernel sums in another order (per-thread, per-tile, replicated atomics). The result
must agree to 1e-13 of the sum of the absolute values of the terms."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, np_ptr, random_pair, random_step
from tests.test_fdm_gpu import device_structs

pytestmark = pytest.mark.gpu
KINDS = {"gt": 0, "i1t": 1, "i2t": 2, "chit": 3}


def gpu_sum(ctx, kind, pairs, temperature, Z):
    import torch
    structs, keep = device_structs(pairs)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    ctx.temp_sum(KINDS[kind], structs, temperature, Z, out)
    torch.cuda.synchronize()
    return float(out.cpu()[0])


def abs_scale(kind, pairs, temperature, Z):
    """sum of |terms|, for the tolerance only"""
    beta = 1.0 / temperature
    s = 0.0
    for p in pairs:
        if not (p.Ki and p.Kj):
            continue
        E1, Ep = p.Ej[:p.Kj, None], p.Ei[None, :p.Ki]
        if kind == "chit":
            x = beta * E1 - beta * Ep
            with np.errstate(divide="ignore", invalid="ignore"):
                st = np.where(np.abs(x) > 1e-8, (np.exp(-beta * Ep) - np.exp(-beta * E1)) / x, np.exp(-beta * Ep))
        else:
            st = beta / (np.exp(beta * E1) + np.exp(beta * Ep)) * np.abs(E1 - Ep) ** KINDS[kind]
        s += abs(p.factor) * np.sum(np.abs(p.op1[:p.Kj, :p.Ki] * p.op2[:p.Kj, :p.Ki]) * np.abs(st)) / Z
    return s


@pytest.mark.parametrize("kind", list(KINDS))
def test_random_steps_match_oracle(ctx, kind):
    for seed in (1, 2):
        pairs = random_step(seed, npairs=16, dmax=200)
        print(f"temp-sum parity: {kind}, seed {seed}", flush=True)
        got = gpu_sum(ctx, kind, pairs, 0.3, 4.0)
        ref = oracle.temp_sum(kind, [p.to_struct(np_ptr) for p in pairs], 0.3, 4.0)
        assert abs(got - ref) <= 1e-13 * abs_scale(kind, pairs, 0.3, 4.0), (got, ref)


def test_conserved_operator_and_large_pair(ctx):
    rng = np.random.default_rng(4)
    m = 700
    E = np.sort(rng.uniform(0, 3, m))
    T = 0.7
    Z = np.exp(-E / T).sum()
    p = Pair(E, E, np.eye(m), np.eye(m), m, m)
    assert gpu_sum(ctx, "chit", [p], T, Z) == pytest.approx(1.0, rel=1e-13)
    assert gpu_sum(ctx, "gt", [p], T, Z) == pytest.approx(1.0 / T / 2.0, rel=1e-13)
    big = [random_pair(rng, 1100, 900, 800, 700, pad=3)]
    got = gpu_sum(ctx, "i2t", big, 0.5, 10.0)
    ref = oracle.temp_sum("i2t", [q.to_struct(np_ptr) for q in big], 0.5, 10.0)
    assert abs(got - ref) <= 1e-13 * abs_scale("i2t", big, 0.5, 10.0)


def test_bad_arguments(ctx):
    import torch
    structs, keep = device_structs(random_step(3, npairs=3, dmax=20))
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for kind, temp, Z in ((4, 1.0, 1.0), (0, 0.0, 1.0), (3, 1.0, -1.0)):
        with pytest.raises(NrgError):
            ctx.temp_sum(kind, structs, temp, Z, out)
    ctx.temp_sum(0, [], 1.0, 1.0, out)
    torch.cuda.synchronize()
    assert float(out.cpu()[0]) == 0.0
