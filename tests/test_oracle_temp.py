"""Pins of the oracle's temperature-dependent sums (Algo_GT<n>: GT, I1T, I2T, c++/algo_FT.hpp:95-134;
Algo_CHIT, c++/algo_FT.hpp:136-170; chit_weight, c++/numerics.hpp:323-336).

Closed forms: a conserved operator (the identity within one subspace, Z the partition function of
that subspace) has chi/beta = <O^2> = 1, through chit_weight's E_n = E_m branch; with the same
identity, GT sums to beta / 2; a single off-diagonal element is checked by hand on both branches;
swapping the two subspaces of a pair flips the sign of I1T and leaves GT, I2T and CHIT unchanged."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, np_ptr


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def run(kind, pairs, temperature, Z):
    return oracle.temp_sum(kind, [p.to_struct(np_ptr) for p in pairs], temperature, Z)


@pytest.mark.parametrize("m", [1, 7, 40])
def test_conserved_operator(m):
    rng = np.random.default_rng(m)
    E = np.sort(rng.uniform(0, 3, m))
    T = 0.7
    Z = np.exp(-E / T).sum()
    p = Pair(E, E, np.eye(m), np.eye(m), m, m)
    assert run("chit", [p], T, Z) == pytest.approx(1.0, rel=1e-13)
    assert run("gt", [p], T, Z) == pytest.approx(1.0 / T / 2.0, rel=1e-13)
    assert run("i1t", [p], T, Z) == 0.0


def test_single_element_by_hand():
    beta = 2.0
    p = Pair([0.1], [0.3], np.array([[1.5]]), np.array([[-2.0]]), 1, 1, factor=0.5)
    w = 1.5 * -2.0 * 0.5
    Z = 3.0
    chi = (math.exp(-0.2) - math.exp(-0.6)) / 0.4 / Z                  # |x| = 0.4 > WEIGHT_TOL
    assert run("chit", [p], 1 / beta, Z) == pytest.approx(w * chi, rel=1e-14)
    g = beta / (math.exp(0.6) + math.exp(0.2)) / Z
    assert run("gt", [p], 1 / beta, Z) == pytest.approx(w * g, rel=1e-14)
    assert run("i1t", [p], 1 / beta, Z) == pytest.approx(w * g * 0.2, rel=1e-13)
    assert run("i2t", [p], 1 / beta, Z) == pytest.approx(w * g * 0.04, rel=1e-13)
    q = Pair([0.1], [0.1 + 1e-10], np.array([[1.5]]), np.array([[-2.0]]), 1, 1, factor=0.5)
    assert run("chit", [q], 1 / beta, Z) == pytest.approx(w * math.exp(-0.2) / Z, rel=1e-14)   # |x| <= WEIGHT_TOL


def test_swapping_the_subspaces():
    rng = np.random.default_rng(3)
    Ei, Ej = np.sort(rng.uniform(0, 2, 5)), np.sort(rng.uniform(0, 2, 4))
    op = rng.standard_normal((4, 5))
    a = Pair(Ei, Ej, op, op, 5, 4)
    b = Pair(Ej, Ei, op.T.copy(), op.T.copy(), 4, 5)
    for kind, s in (("gt", 1), ("i1t", -1), ("i2t", 1), ("chit", 1)):
        assert run(kind, [b], 0.4, 2.0) == pytest.approx(s * run(kind, [a], 0.4, 2.0), rel=1e-12)
    assert run("i1t", [a, b], 0.4, 2.0) == pytest.approx(0.0, abs=1e-14)


def test_discarded_states_do_not_contribute():
    rng = np.random.default_rng(5)
    op = rng.standard_normal((4, 3))
    big = Pair(rng.uniform(0, 1, 3), rng.uniform(0, 1, 4), op, op, 2, 3)
    small = Pair(big.Ei[:2], big.Ej[:3], op[:3, :2], op[:3, :2], 2, 3)
    for kind in ("gt", "i1t", "i2t", "chit"):
        assert run(kind, [big], 0.3, 1.5) == run(kind, [small], 0.3, 1.5)
