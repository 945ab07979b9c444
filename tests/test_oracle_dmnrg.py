"""Pins of the oracle's DM-NRG spectral weights (Algo_DMNRG::calc, c++/algo_DMNRG.hpp:30-52) and of the
N/N+2 window merge (windowfunction / mergeNN2half, c++/spectrum.hpp:91-127).

Expected values come from hand computation (1 x 1 pairs on bin centres) or from properties the
weights must have whatever the basis: the spectral sum rule of a fermion, sum over all weights =
Tr(rho_i d^T d) + Tr(rho_j d d^T) = <{d, d+}> = 1 for an orthogonal d block and unit total trace; the
energy window; and the partition of unity of the windows of consecutive merged steps."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, grid, np_ptr

INF = float("inf")


def run(pairs, g, emin=0.0, emax=INF, sign=-1):
    structs = [p.to_struct(np_ptr) for p in pairs]
    return oracle.dmnrg_bin(structs, g, emin, emax, sign, g.nbins)


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def psd(rng, n, trace):
    a = rng.standard_normal((n, n))
    m = a @ a.T
    return m * (trace / np.trace(m))


def test_single_element_on_bin_centre():
    # weight = (op2 rho_i) op1 + (-sign)(rho_j op1) op2 = b ri a + a rj b for fermions, times the factor
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]),
             factor=0.5)
    pos, neg = run([p], g, sign=-1)
    assert pos[20] == pytest.approx(0.5 * (3.0 * 0.25 * 2.0 + 2.0 * 0.5 * 3.0), rel=1e-15)   # E = 0.1: bin 20
    assert np.count_nonzero(pos) == 1 and not neg.any()
    pos, neg = run([p], g, sign=1)   # bosons: the second term changes sign
    assert pos[20] == pytest.approx(0.5 * (1.5 - 3.0), rel=1e-15)


def test_negative_energy_goes_to_negative_half():
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.3], [0.2], np.array([[1.0]]), np.array([[1.0]]), 1, 1, np.array([[1.0]]), np.array([[0.0]]))
    pos, neg = run([p], g)
    assert not pos.any()
    assert neg[20] == pytest.approx(1.0, rel=1e-15)


def test_discarded_states_do_not_contribute():
    # only the kept x kept block enters (the loops run over diagIp.kept() and diagI1.kept())
    rng = np.random.default_rng(5)
    g = grid(-6.0, 10.0, log10emax=2.0)
    op = rng.standard_normal((4, 3))
    big = Pair(rng.uniform(0, 1, 3), rng.uniform(0, 1, 4), op, op, 2, 3, psd(rng, 2, 0.4), psd(rng, 3, 0.6))
    small = Pair(big.Ei[:2], big.Ej[:3], op[:3, :2], op[:3, :2], 2, 3, big.rho_i, big.rho_j)
    for a, b in zip(run([big], g), run([small], g)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("m", [1, 5, 40])
def test_fermion_sum_rule(m):
    # one fermionic mode times an m-dimensional rest: Ip = mode filled, I1 = mode empty, d = <I1|d|Ip>
    # an arbitrary orthogonal block (basis changes of the two subspaces); rho_i, rho_j positive with
    # Tr rho_i + Tr rho_j = 1. Sum of all weights = Tr(rho_i d^T d) + Tr(rho_j d d^T) = 1.
    rng = np.random.default_rng(m)
    g = grid(-8.0, 10.0, log10emax=2.0)
    d = haar(rng, m)
    ti = rng.uniform(0.2, 0.8)
    p = Pair(rng.uniform(0, 1, m), rng.uniform(0, 1, m), d, d, m, m, psd(rng, m, ti), psd(rng, m, 1 - ti))
    pos, neg = run([p], g, sign=-1)
    assert pos.sum() + neg.sum() == pytest.approx(1.0, abs=1e-13)
    pos, neg = run([p], g, sign=1)    # bosonic sign: Tr rho_i - Tr rho_j
    assert pos.sum() + neg.sum() == pytest.approx(ti - (1 - ti), abs=1e-13)


def test_energy_window():
    # diagonal d and rho: element (m, m) carries rho_i(m,m) + rho_j(m,m) at energy Ej[m] - Ei[m];
    # only |E| in [emin, emax] may contribute
    g = grid(-6.0, 10.0, log10emax=2.0)
    Ei = np.array([0.0, 0.0, 0.5, 0.5, 0.0])
    Ej = np.array([0.001, 0.05, 0.1, 0.55, 3.0])    # |E| = 1e-3, 0.05, 0.4, 0.05, 3
    ri = np.diag([0.1, 0.2, 0.05, 0.15, 0.05])
    rj = np.diag([0.1, 0.1, 0.1, 0.05, 0.1])
    p = Pair(Ei, Ej, np.eye(5), np.eye(5), 5, 5, ri, rj)
    pos, neg = run([p], g, emin=0.01, emax=1.0)
    # m = 0 (|E| = 1e-3) and m = 4 (|E| = 3) lie outside; m = 1 and m = 3 (E = 0.05) are positive,
    # m = 2 (E = -0.4) negative
    assert pos.sum() == pytest.approx(0.3 + 0.2, abs=1e-15)
    assert neg.sum() == pytest.approx(0.15, abs=1e-15)


def test_linear_in_factor_and_additive_over_pairs():
    rng = np.random.default_rng(11)
    g = grid(-6.0, 10.0, log10emax=2.0)
    op1, op2 = rng.standard_normal((6, 5)), rng.standard_normal((6, 5))
    a = Pair(rng.uniform(0, 1, 5), rng.uniform(0, 1, 6), op1, op2, 4, 6, psd(rng, 4, 0.3), psd(rng, 6, 0.7),
             factor=1.0)
    b = Pair(a.Ei, a.Ej, op1, op2, 4, 6, a.rho_i, a.rho_j, factor=2.5)
    one = run([a], g)
    both = run([a, b], g)
    for x, y in zip(one, both):
        assert np.allclose(y, 3.5 * x, rtol=1e-14, atol=1e-16)


# ------------------------------------------------------------------ N/N+2 windows
@pytest.mark.parametrize("nntanh", [0.0, 3.0])
def test_window_endpoints(nntanh):
    Emin, Ex, Emax = 1.0, 2.0, 4.0
    assert oracle.window(Emin, Emin, Ex, Emax, nntanh=nntanh) == 0.0
    assert oracle.window(Emax, Emin, Ex, Emax, nntanh=nntanh) == 0.0
    assert oracle.window(Ex, Emin, Ex, Emax, nntanh=nntanh) == pytest.approx(1.0, abs=1e-15)
    assert oracle.window(Ex * (1 + 1e-12), Emin, Ex, Emax, nntanh=nntanh) == pytest.approx(1.0, abs=1e-10)
    assert oracle.window(0.5, Emin, Ex, Emax, first=True, nntanh=nntanh) == 0.0
    assert oracle.window(3.0, Emin, Ex, Emax, first=True, nntanh=nntanh) == 1.0   # E >= Ex on the first step
    assert oracle.window(1.5, Emin, Ex, Emax, last=True, nntanh=nntanh) == 1.0    # E <= Ex on the last step
    assert oracle.window(10.0, Emin, Ex, Emax, zbw=True, nntanh=nntanh) == 1.0


@pytest.mark.parametrize("nntanh", [0.0, 3.0])
def test_windows_of_consecutive_merged_steps_add_up_to_one(nntanh):
    # merged steps N, N+2, ... have windows [p, p f, p f^2] * scale with the scale falling by f between
    # them (one channel: f = Lambda, two steps of Lambda^(-1/2)); the falling half of one window meets
    # the rising half of the next, so their sum is 1 between the first and the last Ex
    Lam, E0 = 2.0, 1.3
    scales = [Lam ** (-k) for k in range(6)]
    for E in np.exp(np.linspace(np.log(E0 * scales[-1] * Lam * 1.0001), np.log(E0 * Lam * 0.9999), 200)):
        s = sum(oracle.window(E, sc * E0, sc * E0 * Lam, sc * E0 * Lam ** 2, nntanh=nntanh) for sc in scales)
        assert s == pytest.approx(1.0, abs=1e-13)


def test_merge_nn2_weights_bins_by_window():
    g = grid(-2.0, 10.0, log10emax=1.0)
    cs = np.zeros(g.nbins)
    cs[:] = 1.0
    full = np.zeros(g.nbins)
    oracle.merge_nn2(g, full, cs, 0.1, 0.2, 0.4, nn2avg=True)
    e = 10.0 ** (g.log10emin + np.arange(g.nbins) / g.bins_per_decade)
    inside = (e > 0.1) & (e < 0.4)
    assert not full[~inside].any()
    for i in np.nonzero(inside)[0]:
        x = (e[i] - 0.1) / 0.1 if e[i] <= 0.2 else 1.0 - (e[i] - 0.2) / 0.2
        assert full[i] == pytest.approx(0.5 * x, rel=1e-12)
