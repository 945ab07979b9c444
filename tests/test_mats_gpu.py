"""GPU parity of nrg_fdm_mats, nrg_dmnrg_mats and nrg_ft_mats against the oracle (Algo_FDMmats,
c++/algo_FDM.hpp:167-213; Algo_DMNRGmats, c++/algo_DMNRG.hpp:74-100; Algo_FTmats, c++/algo_FT.hpp:59-84).

Tolerance: the kernel adds the elements of a tile in another order than the oracle's loop and forms
1 / (E^2 + w^2) by a refined reciprocal instead of the complex division. Every term is bounded by
|w_e| / (pi T) (|i w_n - E| >= w_n >= pi T for n >= 1, and for fermions at n = 0), so every
frequency must agree to 1e-12 of S = sum_e |w_e| / (pi T); the bosonic n = 0 value, where terms are
w / E, to 1e-12 of S pi T / min |E|."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError
from tests.fdm_cases import Pair, grid, np_ptr, random_pair, random_step
from tests.test_fdm_gpu import device_structs

pytestmark = pytest.mark.gpu
T = 0.5
INF = float("inf")


def gpu_mats(ctx, algo, pairs, mats, sign, wnf=1.0, Z=1.0, G=None):
    import torch
    structs, keep = device_structs(pairs)
    if G is None:
        G = torch.zeros(2 * max(mats, 1), dtype=torch.float64, device="cuda")
    ctx.mats(algo, structs, mats, T, sign, G, wnf=wnf, Z=Z)
    torch.cuda.synchronize()
    return G.cpu().numpy()[:2 * mats].view(np.complex128)


def total_abs_weight(algo, pairs, wnf, Z):
    """sum_e |w_e| from the oracle's binned form of the same algorithm on absolute values."""
    ab = [Pair(p.Ei, p.Ej, np.abs(p.op1[:, :p.di]), np.abs(p.op2[:, :p.di]), p.Ki, p.Kj,
               None if p.rho_i is None else np.abs(p.rho_i), None if p.rho_j is None else np.abs(p.rho_j),
               abs(p.factor)) for p in pairs]
    s = [p.to_struct(np_ptr) for p in ab]
    g = grid(-12.0, 1.0, log10emax=4.0)
    if algo == "fdm":
        out = oracle.fdm_bin(s, g, T, wnf, -1, g.nbins)
    elif algo == "dmnrg":
        out = oracle.dmnrg_bin(s, g, 0.0, INF, -1, g.nbins)
    else:
        out = oracle.ft_bin(s, g, T, Z, -1, g.nbins)
    return sum(float(np.abs(o).sum()) for o in out)


def min_energy(pairs):
    m = INF
    for p in pairs:
        if p.di and p.dj:
            m = min(m, float(np.min(np.abs(p.Ej[:, None] - p.Ei[None, :]))))
    return m


def check(ctx, algo, pairs, mats, sign, wnf=1.0, Z=1.0):
    print(f"mats parity: {algo}, {len(pairs)} pairs, mats {mats}, sign {sign}", flush=True)
    got = gpu_mats(ctx, algo, pairs, mats, sign, wnf, Z)
    ref = oracle.mats(algo, [p.to_struct(np_ptr) for p in pairs], mats, T, sign, wnf=wnf, Z=Z)
    S = total_abs_weight(algo, pairs, wnf, Z) / (math.pi * T)
    err = np.abs(got - ref)
    n0 = 1 if sign == 1 else 0
    if mats > n0:
        assert np.max(err[n0:]) <= 1e-12 * S, (np.max(err[n0:]), S)
    if sign == 1 and mats > 0:
        assert err[0] <= 1e-12 * S * math.pi * T / min_energy(pairs), (err[0], S)


def with_rho(pairs, which):
    """Give every pair the density matrices its algorithm reads."""
    rng = np.random.default_rng(77)
    for p in pairs:
        if which == "dmnrg" and p.Ki and p.Kj:
            if p.rho_i is None or p.rho_i.shape[0] < p.Ki:
                p.rho_i = rng.standard_normal((p.Ki, p.Ki)) / p.Ki
            if p.rho_j is None or p.rho_j.shape[0] < p.Kj:
                p.rho_j = rng.standard_normal((p.Kj, p.Kj)) / p.Kj
    return pairs


def test_hand_values(ctx):
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]),
             factor=0.5)
    G = gpu_mats(ctx, "dmnrg", [p], 3, -1)
    for n in range(3):
        assert G[n] == pytest.approx(0.5 * 4.5 / (1j * T * math.pi * (2 * n + 1) - 0.1), rel=1e-14)
    q = Pair([0.2], [0.2], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]))
    G = gpu_mats(ctx, "dmnrg", [q], 2, 1)
    assert G[0] == pytest.approx(-1.5 / T, rel=1e-14)       # the bosonic limit -w_A / T at E = 0
    assert G[1] == pytest.approx(-1.5 / (1j * 2 * math.pi * T), rel=1e-14)


@pytest.mark.parametrize("algo", ["fdm", "dmnrg", "ft"])
@pytest.mark.parametrize("sign", [-1, 1])
def test_random_steps_match_oracle(ctx, algo, sign):
    pairs = with_rho(random_step(3, npairs=12, dmax=120), algo)
    for mats in (1, 50, 300):
        check(ctx, algo, pairs, mats, sign, wnf=0.37, Z=5.0)


def test_many_frequencies_and_a_large_pair(ctx):
    rng = np.random.default_rng(12)
    pairs = [random_pair(rng, 700, 500, 300, 200, pad=3), random_pair(rng, 90, 130, 40, 130)]
    check(ctx, "fdm", pairs, 1100, -1, wnf=0.5)          # several frequency chunks
    check(ctx, "ft", [random_pair(rng, 400, 300, 400, 300)], 200, 1, Z=30.0)


def test_high_frequency_tail(ctx):
    rng = np.random.default_rng(3)
    m = 200
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    Ei, Ej = np.sort(rng.uniform(0, 1, m)), np.sort(rng.uniform(0, 1, m))
    Z = np.exp(-Ei / T).sum() + np.exp(-Ej / T).sum()
    G = gpu_mats(ctx, "ft", [Pair(Ei, Ej, d, d, m, m)], 2000, -1, Z=Z)
    w = T * math.pi * (2 * 1999 + 1)
    assert abs(1j * w * G[1999] - 1.0) <= 2.0 / w


def test_accumulates_and_bad_arguments(ctx):
    import torch
    pairs = with_rho(random_step(9, npairs=4, dmax=60), "dmnrg")
    G = torch.zeros(2 * 40, dtype=torch.float64, device="cuda")
    once = gpu_mats(ctx, "dmnrg", pairs, 40, -1, G=G).copy()
    twice = gpu_mats(ctx, "dmnrg", pairs, 40, -1, G=G)
    np.testing.assert_allclose(twice, 2 * once, rtol=1e-13, atol=1e-13 * np.max(np.abs(once)))
    structs, keep = device_structs(pairs)
    for kw in (dict(mats=-1), dict(mats=40000), dict(sign=0), dict(T=0.0)):
        a = dict(mats=40, sign=-1, T=T)
        a.update(kw)
        with pytest.raises(NrgError):
            ctx.mats("dmnrg", structs, a["mats"], a["T"], a["sign"], G)
    with pytest.raises(NrgError):
        ctx.mats("ft", structs, 40, T, -1, G, Z=0.0)
    bad = [type(s).from_buffer_copy(s) for s in structs]
    for b in bad:
        b.rho_i = None
    if any(p.Ki and p.Kj for p in pairs):
        with pytest.raises(NrgError):
            ctx.mats("dmnrg", bad, 40, T, -1, G)
    ctx.mats("ft", bad, 40, T, -1, G, Z=2.0)     # FT reads no density matrix
    ctx.mats("fdm", structs, 0, T, -1, G)         # nothing to add
    torch.cuda.synchronize()
