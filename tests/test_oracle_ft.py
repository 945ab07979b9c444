"""Pins of the oracle's FT spectral weights (Algo_FT::calc, c++/algo_FT.hpp:28-44).

Expected values come from hand computation (a 1 x 1 pair on a bin centre, both signs) and from the
sum rule of an orthogonal d block: sum_{i,j} d(j,i)^2 ((-sign) exp(-E_j/T) + exp(-E_i/T)) / Z =
(Z_i - sign Z_j) / Z, which is 1 for fermions when Z is the partition function of both subspaces."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, grid, np_ptr

T = 0.5


def run(pairs, g, sign=-1, Z=1.0):
    return oracle.ft_bin([p.to_struct(np_ptr) for p in pairs], g, T, Z, sign, g.nbins)


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def test_single_element_both_signs():
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, factor=0.5)
    pos, neg = run([p], g, sign=-1, Z=4.0)
    assert pos[20] == pytest.approx(6.0 * (math.exp(-0.3 / T) + math.exp(-0.2 / T)) / 4.0 * 0.5, rel=1e-15)
    assert np.count_nonzero(pos) == 1 and not neg.any()
    pos, neg = run([p], g, sign=1, Z=4.0)
    assert pos[20] == pytest.approx(6.0 * (math.exp(-0.2 / T) - math.exp(-0.3 / T)) / 4.0 * 0.5, rel=1e-14)
    # the reversed pair lies at -0.1
    q = Pair([0.3], [0.2], np.array([[2.0]]), np.array([[3.0]]), 1, 1)
    pos, neg = run([q], g, Z=4.0)
    assert not pos.any() and neg[20] == pytest.approx(6.0 * (math.exp(-0.2 / T) + math.exp(-0.3 / T)) / 4.0, rel=1e-15)


@pytest.mark.parametrize("m", [1, 5, 30])
def test_sum_rule(m):
    rng = np.random.default_rng(m)
    q, r = np.linalg.qr(rng.standard_normal((m, m)))
    d = q * np.sign(np.diag(r))
    Ei, Ej = np.sort(rng.uniform(0, 2, m)), np.sort(rng.uniform(0, 2, m))
    Zi, Zj = np.exp(-Ei / T).sum(), np.exp(-Ej / T).sum()
    g = grid(-8.0, 10.0, log10emax=2.0)
    pos, neg = run([Pair(Ei, Ej, d, d, m, m)], g, sign=-1, Z=Zi + Zj)
    assert pos.sum() + neg.sum() == pytest.approx(1.0, abs=1e-13)
    pos, neg = run([Pair(Ei, Ej, d, d, m, m)], g, sign=1, Z=Zi + Zj)
    assert pos.sum() + neg.sum() == pytest.approx((Zi - Zj) / (Zi + Zj), abs=1e-13)


def test_discarded_states_do_not_contribute():
    rng = np.random.default_rng(5)
    g = grid(-6.0, 10.0, log10emax=2.0)
    op1, op2 = rng.standard_normal((4, 3)), rng.standard_normal((4, 3))
    big = Pair(rng.uniform(0, 1, 3), rng.uniform(0, 1, 4), op1, op2, 2, 3)
    small = Pair(big.Ei[:2], big.Ej[:3], op1[:3, :2], op2[:3, :2], 2, 3)
    for a, b in zip(run([big], g, Z=2.0), run([small], g, Z=2.0)):
        assert np.array_equal(a, b)
