"""Pins of the oracle's CFS spectral weights (Algo_CFS::calc = Algo_CFSgt + Algo_CFSls,
c++/algo_CFS.hpp:29-56, 79-106, 129-133; prod_fit_left / prod_adj_fit_left, c++/numerics.hpp:283-316).

Expected values come from hand computation (2 x 2 pairs with non-symmetric density matrices, so a
transposed product or a swapped operator changes the answer) and from sum rules that hold in any
basis: for an orthogonal d block, sum_{j, i} d(j, i) (d rho_i)(j, i) = Tr rho_i over the kept i when
every j is discarded (Eq. 16), sum (rho_j^T d)(j, i) d(j, i) = Tr rho_j over the kept j when every i is
discarded (Eq. 15), and the last-step weights (Eq. 11) add up to <{d, d+}> = 1 with Z the partition
function of both subspaces."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, grid, np_ptr

T = 0.5


def run(pairs, g, last, sign=-1, T_=T, Z=1.0):
    structs = [p.to_struct(np_ptr) for p in pairs]
    return oracle.cfs_bin(structs, g, T_, Z, last, sign, g.nbins)


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


def test_last_step_single_element():
    # gt: op1 op2 exp(-E_i/T) / Z, ls: op1 op2 exp(-E_j/T) (-sign) / Z, both at E_j - E_i = 0.1 (bin 20)
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.2], [0.3], np.array([[2.0]]), np.array([[3.0]]), 1, 1, factor=0.5)
    gp, gn, lp, ln = run([p], g, last=True, Z=4.0)
    assert gp[20] == pytest.approx(6.0 * math.exp(-0.2 / T) / 4.0 * 0.5, rel=1e-15)
    assert lp[20] == pytest.approx(6.0 * math.exp(-0.3 / T) / 4.0 * 0.5, rel=1e-15)
    assert np.count_nonzero(gp) == 1 and np.count_nonzero(lp) == 1 and not gn.any() and not ln.any()
    gp2, _, lp2, _ = run([p], g, last=True, sign=1, Z=4.0)   # bosons: the lesser half changes sign
    assert gp2[20] == gp[20] and lp2[20] == -lp[20]


def hand_pair():
    # i: Ip (E 0.125 kept, 0.5 discarded), j: I1 (E 0.25 kept, 1.125 discarded); op(j, i)
    op1 = np.array([[1.0, 2.0], [3.0, 4.0]])
    op2 = np.array([[5.0, 6.0], [7.0, 8.0]])
    rho_i = np.array([[0.5, 0.25], [0.125, 0.0625]])
    rho_j = np.array([[0.3, 0.2], [0.1, 0.4]])
    return Pair([0.125, 0.5], [0.25, 1.125], op1, op2, 1, 1, rho_i, rho_j)


def test_below_last_step_hand_values():
    g = grid(-3.0, 10.0, log10emax=2.0)
    gp, gn, lp, ln = run([hand_pair()], g, last=False)
    # Eq. (16), j = 1 discarded, i = 0 kept: op1(1,0) (op2 rho_i)(1,0) (-sign)
    #   = 3 (7 * 0.5 + 8 * 0.125) = 13.5 at E = 1.125 - 0.125 = 1 (bin 30)
    assert lp[30] == 13.5 and np.count_nonzero(lp) == 1 and not ln.any()
    # Eq. (15), j = 0 kept, i = 1 discarded: (op1^T rho_j)(1,0) op2(0,1) = (2 * 0.3 + 4 * 0.1) * 6 = 6
    #   at E = 0.25 - 0.5 = -0.25
    assert not gp.any() and gn.sum() == pytest.approx(6.0, rel=1e-15)
    nz = np.nonzero(gn)[0]
    x = (math.log10(0.25) + 3.0) * 10.0
    assert list(nz) == [math.floor(x), math.floor(x) + 1]
    assert gn[nz[1]] / 6.0 == pytest.approx(x - math.floor(x), rel=1e-12)
    _, _, lp_b, _ = run([hand_pair()], g, last=False, sign=1)
    assert lp_b[30] == -13.5


def test_last_step_hand_values():
    # only kept x kept (0, 0): op1 op2 = 5 at E = 0.25 - 0.125 = 0.125
    g = grid(-3.0, 10.0, log10emax=2.0)
    gp, gn, lp, ln = run([hand_pair()], g, last=True, Z=2.0)
    assert gp.sum() == pytest.approx(5.0 * math.exp(-0.125 / T) / 2.0, rel=1e-15)
    assert lp.sum() == pytest.approx(5.0 * math.exp(-0.25 / T) / 2.0, rel=1e-15)
    assert not gn.any() and not ln.any()


def test_unread_regions():
    # with density matrices over the kept states (their size in the reference), nothing in kept x kept
    # or discarded x discarded is read below the last step; at the last step nothing outside kept x kept
    rng = np.random.default_rng(4)
    g = grid(-6.0, 10.0, log10emax=2.0)
    di, dj, Ki, Kj = 7, 6, 3, 2
    op1, op2 = rng.standard_normal((dj, di)), rng.standard_normal((dj, di))
    base = Pair(np.sort(rng.uniform(0, 1, di)), np.sort(rng.uniform(0, 1, dj)), op1, op2, Ki, Kj,
                rng.standard_normal((Ki, Ki)), rng.standard_normal((Kj, Kj)))
    ref_nl, ref_l = run([base], g, False), run([base], g, True, Z=3.0)
    o1, o2 = op1.copy(), op2.copy()
    o1[:Kj, :Ki] = o2[:Kj, :Ki] = 99.0                 # kept x kept
    o1[Kj:, Ki:] = o2[Kj:, Ki:] = -77.0                # discarded x discarded
    changed = Pair(base.Ei, base.Ej, o1, o2, Ki, Kj, base.rho_i, base.rho_j)
    for a, b in zip(ref_nl, run([changed], g, False)):
        assert np.array_equal(a, b)
    o1, o2 = op1.copy(), op2.copy()
    o1[Kj:, :] = o2[Kj:, :] = 55.0
    o1[:, Ki:] = o2[:, Ki:] = 33.0
    changed = Pair(base.Ei, base.Ej, o1, o2, Ki, Kj, None, None)
    for a, b in zip(ref_l, run([changed], g, True, Z=3.0)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("m", [1, 6, 40])
def test_sum_rules_below_last_step(m):
    rng = np.random.default_rng(m)
    g = grid(-8.0, 10.0, log10emax=2.0)
    d = haar(rng, m)
    ri, rj = psd(rng, m, 0.7), psd(rng, m, 0.3)
    E = (np.sort(rng.uniform(0, 1, m)), np.sort(rng.uniform(0, 1, m)))
    # every j discarded, every i kept: the lesser half carries Tr rho_i (fermions), the greater nothing
    gp, gn, lp, ln = run([Pair(*E, d, d, m, 0, ri, None)], g, False)
    assert lp.sum() + ln.sum() == pytest.approx(0.7, abs=1e-13) and not gp.any() and not gn.any()
    # every i discarded, every j kept: the greater half carries Tr rho_j
    gp, gn, lp, ln = run([Pair(*E, d, d, 0, m, None, rj)], g, False)
    assert gp.sum() + gn.sum() == pytest.approx(0.3, abs=1e-13) and not lp.any() and not ln.any()
    # the first k states of i kept: the diagonal of rho_i over them
    k = max(m // 2, 1)
    gp, gn, lp, ln = run([Pair(*E, d, d, k, 0, ri, None)], g, False)
    assert lp.sum() + ln.sum() == pytest.approx(np.trace(ri[:k, :k]), abs=1e-13)


@pytest.mark.parametrize("m", [1, 5, 30])
def test_last_step_sum_rule(m):
    rng = np.random.default_rng(10 + m)
    g = grid(-8.0, 10.0, log10emax=2.0)
    d = haar(rng, m)
    Ei, Ej = np.sort(rng.uniform(0, 2, m)), np.sort(rng.uniform(0, 2, m))
    Z = np.exp(-Ei / T).sum() + np.exp(-Ej / T).sum()
    out = run([Pair(Ei, Ej, d, d, m, m)], g, True, Z=Z)
    assert sum(o.sum() for o in out) == pytest.approx(1.0, abs=1e-13)
    gt = out[0].sum() + out[1].sum()
    assert gt == pytest.approx(np.exp(-Ei / T).sum() / Z, abs=1e-13)   # the greater half: the occupied i side


def test_linear_in_factor_and_additive_over_pairs():
    rng = np.random.default_rng(11)
    g = grid(-6.0, 10.0, log10emax=2.0)
    op1, op2 = rng.standard_normal((6, 5)), rng.standard_normal((6, 5))
    a = Pair(rng.uniform(0, 1, 5), rng.uniform(0, 1, 6), op1, op2, 3, 2, psd(rng, 4, 0.3), psd(rng, 5, 0.7))
    b = Pair(a.Ei, a.Ej, op1, op2, 3, 2, a.rho_i, a.rho_j, factor=2.5)
    for last in (False, True):
        one = run([a], g, last, Z=2.0)
        both = run([a, b], g, last, Z=2.0)
        for x, y in zip(one, both):
            assert np.allclose(y, 3.5 * x, rtol=1e-14, atol=1e-16)
