"""Pins of the oracle's Matsubara-axis Green's functions (Algo_FDMmats, c++/algo_FDM.hpp:167-213;
Algo_DMNRGmats, c++/algo_DMNRG.hpp:74-100; Algo_FTmats, c++/algo_FT.hpp:59-84; ww, c++/matsubara.hpp).

What fixes them independently of the restatement:
  - hand values of single elements, fermionic and bosonic, including the bosonic n = 0 limit -w_A/T
    at E = 0;
  - the spectral representation: with every excitation energy on a bin centre, G(i w_n) equals
    sum_b A_b / (i w_n - e_b) over the bins of the real-frequency algorithm of the same name (FDM,
    DM-NRG, FT), which the oracle's binned functions compute from their own pinned terms;
  - the high-frequency tail i w_n G(i w_n) -> total weight = 1 for a fermionic FT step with an
    orthogonal d (the same sum rule as the real-frequency FT)."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, grid, np_ptr

T = 0.5
INF = float("inf")


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def S(pairs):
    """Descriptors pointing into the pairs' arrays: the caller keeps the pairs alive."""
    return [p.to_struct(np_ptr) for p in pairs]


def w_n(n, sign):
    return T * math.pi * (2 * n + (0 if sign == 1 else 1))


def test_ft_single_element_both_statistics():
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, factor=0.5)
    G = oracle.mats("ft", S([p]), 4, T, -1, Z=4.0)
    num = 6.0 * (math.exp(-0.3 / T) + math.exp(-0.2 / T)) / 4.0
    for n in range(4):
        assert G[n] == pytest.approx(0.5 * num / (1j * w_n(n, -1) - 0.1), rel=1e-14)
    G = oracle.mats("ft", S([p]), 4, T, 1, Z=4.0)
    num = 6.0 * (math.exp(-0.2 / T) - math.exp(-0.3 / T)) / 4.0
    assert G[0] == pytest.approx(0.5 * num / (-0.1), rel=1e-14)          # w_0 = 0 for bosons
    assert G[2] == pytest.approx(0.5 * num / (1j * w_n(2, 1) - 0.1), rel=1e-14)
    # bosonic E = 0: n = 0 takes the limit -exp(-E_1/T) / (Z T) times op1 op2
    q = Pair([0.25], [0.25], np.array([[2.0]]), np.array([[3.0]]), 1, 1)
    G = oracle.mats("ft", S([q]), 2, T, 1, Z=4.0)
    assert G[0] == pytest.approx(-6.0 * math.exp(-0.25 / T) / (4.0 * T), rel=1e-14)
    assert G[1] == 0.0    # (-sign) e^{-E/T} + e^{-E/T} = 0 for bosons at E = 0


def test_dmnrg_single_element():
    # w_A = (op2 rho_i) op1, w_B = (rho_j op1) op2; G = f (w_A - sign w_B) / (i w_n - E)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]),
             factor=0.5)
    G = oracle.mats("dmnrg", S([p]), 3, T, -1)
    for n in range(3):
        assert G[n] == pytest.approx(0.5 * (1.5 + 3.0) / (1j * w_n(n, -1) - 0.1), rel=1e-14)
    q = Pair([0.2], [0.2], np.array([[2.0]]), np.array([[3.0]]), 1, 1, np.array([[0.25]]), np.array([[0.5]]))
    G = oracle.mats("dmnrg", S([q]), 2, T, 1)
    assert G[0] == pytest.approx(-1.5 / T, rel=1e-14)                    # the limit -w_A / T
    assert G[1] == pytest.approx((1.5 - 3.0) / (1j * w_n(1, 1)), rel=1e-14)


def test_fdm_terms_by_hand():
    # 2 x 2 pair, K = 1 on both sides: (i, j) = (1, 1) term1, (1, 0) term2, (0, 1) term3
    op1 = np.array([[1.0, 2.0], [3.0, 4.0]])
    op2 = np.array([[5.0, 6.0], [7.0, 8.0]])
    ri = np.array([[0.5, 0.25], [0.125, 0.0625]])
    rj = np.array([[0.3, 0.2], [0.1, 0.4]])
    Ei, Ej = [0.125, 0.5], [0.25, 1.125]
    wnf = 0.75
    p = Pair(Ei, Ej, op1, op2, 1, 1, ri, rj)
    G = oracle.mats("fdm", S([p]), 3, T, -1, wnf=wnf)
    b = lambda e: math.exp(-e / T)
    t1 = (4.0 * 8.0 * wnf * b(0.5) + 4.0 * 8.0 * wnf * b(1.125), 1.125 - 0.5)
    # term2: op1(0,1) (rho_j op2)(0,1), (rho_j op2)(0,1) = 0.3 * 6 + 0.2 * 8 = 3.4
    t2 = (2.0 * 6.0 * wnf * b(0.5) + 2.0 * 3.4, 0.25 - 0.5)
    # term3: op1(1,0) (op2 rho_i)(1,0), (op2 rho_i)(1,0) = 7 * 0.5 + 8 * 0.125 = 4.5
    t3 = (3.0 * 4.5 + 3.0 * 7.0 * wnf * b(1.125), 1.125 - 0.125)
    for n in range(3):
        want = sum(w / (1j * w_n(n, -1) - e) for w, e in (t1, t2, t3))
        assert G[n] == pytest.approx(want, rel=1e-13)


def centre_pairs(with_rho):
    """Pairs whose excitation energies E_j - E_i all sit on centres of a 1-bin-per-decade grid."""
    rng = np.random.default_rng(8)
    a = Pair([0.0, 0.0], [0.1, 1.0, 100.0], rng.standard_normal((3, 2)), rng.standard_normal((3, 2)), 1, 2,
             rng.standard_normal((1, 1)) if with_rho else None, rng.standard_normal((2, 2)) if with_rho else None,
             factor=0.7)
    b = Pair([0.1, 10.0, 1.0], [0.0, 0.0], rng.standard_normal((2, 3)), rng.standard_normal((2, 3)), 2, 1,
             rng.standard_normal((2, 2)) if with_rho else None, rng.standard_normal((1, 1)) if with_rho else None,
             factor=1.3)
    return [a, b]


def from_bins(g, halves, n, sign):
    e = 10.0 ** (g.log10emin + np.arange(g.nbins) / g.bins_per_decade)
    z = 1j * w_n(n, sign)
    out = 0.0
    for pos, neg in halves:
        out += np.sum(pos / (z - e)) + np.sum(neg / (z + e))
    return out


@pytest.mark.parametrize("sign", [-1, 1])
def test_spectral_representation_of_the_binned_algorithms(sign):
    g = grid(-3.0, 1.0, log10emax=3.0)
    pairs = centre_pairs(True)
    mats = 6
    # FDM: gt + ls halves
    wnf = 0.6
    b = oracle.fdm_bin(S(pairs), g, T, wnf, sign, g.nbins)
    G = oracle.mats("fdm", S(pairs), mats, T, sign, wnf=wnf)
    for n in range(1 if sign == 1 else 0, mats):
        assert G[n] == pytest.approx(from_bins(g, [(b[0], b[1]), (b[2], b[3])], n, sign), rel=1e-12, abs=1e-14)
    # DM-NRG with the window wide open
    b = oracle.dmnrg_bin(S(pairs), g, 0.0, INF, sign, g.nbins)
    G = oracle.mats("dmnrg", S(pairs), mats, T, sign)
    for n in range(1 if sign == 1 else 0, mats):
        assert G[n] == pytest.approx(from_bins(g, [(b[0], b[1])], n, sign), rel=1e-12, abs=1e-14)
    # FT
    b = oracle.ft_bin(S(pairs), g, T, 3.0, sign, g.nbins)
    G = oracle.mats("ft", S(pairs), mats, T, sign, Z=3.0)
    for n in range(1 if sign == 1 else 0, mats):
        assert G[n] == pytest.approx(from_bins(g, [(b[0], b[1])], n, sign), rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("m", [3, 20])
def test_high_frequency_tail(m):
    rng = np.random.default_rng(m)
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    Ei, Ej = np.sort(rng.uniform(0, 1, m)), np.sort(rng.uniform(0, 1, m))
    Z = np.exp(-Ei / T).sum() + np.exp(-Ej / T).sum()
    mats = 2000
    pair = Pair(Ei, Ej, d, d, m, m)   # held: the descriptors point into its arrays
    G = oracle.mats("ft", S([pair]), mats, T, -1, Z=Z)
    n = mats - 1
    tail = 1j * w_n(n, -1) * G[n]
    assert abs(tail - 1.0) <= 2.0 / w_n(n, -1)     # the next order is sum w E / (i w_n), |E| < 1
    assert G[0].imag < 0                            # the retarded-like sign convention 1/(i w - E)
