"""Pins of the oracle's FDM weights + binning (Algo_FDMgt/ls c++/algo_FDM.hpp:30-60,84-114,
ChainBinning c++/spectrum.hpp:19-28, Bins::add/add_std c++/bins.hpp:103-132).

Every expected value below is computed by hand (small integer matrices, energies placed on bin
centres) or follows from a property of the binning (conservation, first moment, sign split), not by
restating the oracle's loops."""
import math

import numpy as np
import pytest

import oracle
from tests.fdm_cases import Pair, grid, np_ptr, random_pair

T = 0.1
WNF = 0.75


def run(pairs, g, sign=-1, T=T, wnf=WNF):
    structs = [p.to_struct(np_ptr) for p in pairs]
    return oracle.fdm_bin(structs, g, T, wnf, sign, g.nbins)


def bin_of(E, g):
    """Position of energy E on the standard grid: integer part and remainder of
    x = (log10 E - log10emin) * bins_per_decade."""
    x = (math.log10(E) - g.log10emin) * g.bins_per_decade
    return int(math.floor(x)), x - math.floor(x)


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def test_single_peak_on_bin_centre():
    # one discarded state on each side: gt weight a b exp(-Ei/T) wnf, ls weight a b (-sign) exp(-Ej/T) wnf
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.0], [0.1], np.array([[2.0]]), np.array([[3.0]]), 0, 0, factor=0.5)
    gtp, gtn, lsp, lsn = run([p], g, sign=-1)
    b, rem = bin_of(0.1, g)
    assert (b, rem) == (20, 0.0)
    assert gtp[20] == pytest.approx(0.5 * 6.0 * WNF, rel=1e-15)
    assert lsp[20] == pytest.approx(0.5 * 6.0 * math.exp(-1.0) * WNF, rel=1e-15)
    for a in (gtp, lsp):
        assert np.count_nonzero(a) == 1
    assert not gtn.any() and not lsn.any()


def test_negative_energy_goes_to_neg_arrays_with_abs_energy():
    g = grid(-3.0, 10.0, log10emax=2.0)
    p = Pair([0.1], [0.0], np.array([[1.0]]), np.array([[1.0]]), 0, 0)
    gtp, gtn, lsp, lsn = run([p], g, sign=1)
    assert not gtp.any() and not lsp.any()
    assert gtn[20] == pytest.approx(math.exp(-1.0) * WNF, rel=1e-15)
    assert lsn[20] == pytest.approx(-1.0 * WNF, rel=1e-15)   # bosons: -sign = -1, exp(-0/T) = 1


def test_split_between_bins_conserves_weight_and_first_moment():
    g = grid(-4.0, 7.0, log10emax=1.0)
    for E in (3.7e-3, 0.0123, 0.5, 1.9e-4):
        p = Pair([0.0], [E], np.array([[1.5]]), np.array([[-0.25]]), 0, 0)
        gtp, _, _, _ = run([p], g)
        w = 1.5 * -0.25 * WNF
        b, rem = bin_of(E, g)
        nz = np.nonzero(gtp)[0]
        assert set(nz) <= {b, b + 1}
        assert gtp.sum() == pytest.approx(w, rel=1e-15)
        # the unbiased split keeps the log-position: sum_k k w_k = x w
        assert (np.arange(g.nbins) * gtp).sum() == pytest.approx((b + rem) * w, rel=1e-13)
        assert gtp[b] == pytest.approx((1 - rem) * w, rel=1e-14, abs=1e-300)


def test_clamps_zero_low_and_high_energies():
    g = grid(-3.0, 10.0, log10emax=1.0)
    # energy 0 (degenerate states), below the grid, and above the grid
    for Ei, Ej, where in ((0.3, 0.3, 0), (0.0, 1e-13, 0), (0.0, 2e-6, 0), (0.0, 50.0, g.nbins - 1)):
        p = Pair([Ei], [Ej], np.array([[1.0]]), np.array([[1.0]]), 0, 0)
        gtp, gtn, _, _ = run([p], g)
        assert np.nonzero(gtp)[0].tolist() == [where]
        assert gtp[where] == pytest.approx(math.exp(-Ei / T) * WNF, rel=1e-15)
        assert not gtn.any()


def test_discard_immediately():
    # |w| < discard * E is dropped (c++/bins.hpp:105)
    g = grid(-3.0, 10.0, log10emax=1.0, discard=1e-3)
    E = 0.1
    keep = Pair([0.0], [E], np.array([[1.0]]), np.array([[2e-4 / WNF]]), 0, 0)   # w = 2e-4 >= 1e-4
    drop = Pair([0.0], [E], np.array([[1.0]]), np.array([[0.5e-4 / WNF]]), 0, 0)  # w = 5e-5 < 1e-4
    gtp, _, _, _ = run([keep], g)
    assert gtp.sum() == pytest.approx(2e-4, rel=1e-14)
    gtp, _, _, _ = run([drop], g)
    assert not gtp.any()


def test_fermion_boson_sign_only_flips_lesser():
    rng = np.random.default_rng(5)
    pairs = [random_pair(rng, 7, 9, 3, 4) for _ in range(3)]
    g = grid()
    f = run(pairs, g, sign=-1)
    b = run(pairs, g, sign=1)
    np.testing.assert_array_equal(f[0], b[0])
    np.testing.assert_array_equal(f[1], b[1])
    np.testing.assert_array_equal(f[2], -b[2])
    np.testing.assert_array_equal(f[3], -b[3])


def test_greater_kept_discarded_uses_op2_times_rho_i():
    # i kept (K_i = dim_i = 2), j discarded (dim_j = 1, K_j = 0): gt weight op1(0,i) (op2 rho_i)(0,i)
    op1 = np.array([[1.0, 2.0]])
    op2 = np.array([[3.0, 5.0]])
    rho_i = np.array([[7.0, 11.0], [13.0, 17.0]])   # op2 rho = [86, 118]; op2 rho^T would be [76, 124]
    p = Pair([0.0, 0.0], [0.1], op1, op2, 2, 0, rho_i=rho_i, factor=2.0)
    g = grid(-3.0, 10.0, log10emax=1.0)
    gtp, gtn, lsp, lsn = run([p], g, sign=-1)
    assert gtp[20] == 2.0 * (1 * 86 + 2 * 118)
    # ls: K_i x D_j uses op1 op2 (-sign) exp(-Ej/T) wnf = (3 + 10) e^-1 wnf
    assert lsp[20] == pytest.approx(2.0 * 13.0 * math.exp(-1.0) * WNF, rel=1e-15)
    assert np.count_nonzero(gtp) == 1 and np.count_nonzero(lsp) == 1


def test_lesser_discarded_kept_uses_rho_j_times_op2():
    # i discarded (dim_i = 1, K_i = 0), j kept (K_j = dim_j = 2): ls weight op1(j,0) (rho_j op2)(j,0) (-sign)
    op1 = np.array([[1.0], [2.0]])
    op2 = np.array([[3.0], [5.0]])
    rho_j = np.array([[7.0, 11.0], [13.0, 17.0]])   # rho op2 = [76, 124]
    p = Pair([0.0], [0.1, 0.1], op1, op2, 0, 2, rho_j=rho_j)
    g = grid(-3.0, 10.0, log10emax=1.0)
    gtp, _, lsp, _ = run([p], g, sign=1)
    assert lsp[20] == -(1 * 76 + 2 * 124)
    # gt: D_i x K_j uses op1 op2 exp(-Ei/T) wnf = (3 + 10) wnf
    assert gtp[20] == pytest.approx(13.0 * WNF, rel=1e-15)


def test_prod_fit_size_adaptation():
    # rho larger than the subspace (prod_fit_right): rows/cols beyond dim are never used
    op1 = np.array([[1.0, 2.0]])
    op2 = np.array([[3.0, 5.0]])
    big = np.full((3, 3), 1e6)
    big[:2, :2] = [[7.0, 11.0], [13.0, 17.0]]
    g = grid(-3.0, 10.0, log10emax=1.0)
    p = Pair([0.0, 0.0], [0.1], op1, op2, 2, 0, rho_i=big)
    gtp, _, _, _ = run([p], g)
    assert gtp[20] == 1 * 86 + 2 * 118
    # rho covering only the kept states (prod_fit_left): the discarded column of op2 is not contracted
    op1 = np.array([[1.0, 2.0, 4.0]])
    op2 = np.array([[3.0, 5.0, 1e6]])
    p = Pair([0.0, 0.0, 0.0], [0.1], op1, op2, 2, 0, rho_i=big[:2, :2].copy())
    gtp, _, _, _ = run([p], g)
    # K_i x D_j: 1*86 + 2*118; D_i x D_j (i = 2): 4 * 1e6 * wnf
    assert gtp[20] == pytest.approx(1 * 86 + 2 * 118 + 4e6 * WNF, rel=1e-15)


def test_boltzmann_density_matrix_reproduces_all_discarded_result():
    # with rho = diag(exp(-E/T) wnf) on the kept states, the K x D (gt) and D x K (ls) terms equal the
    # D formulas, so FDM(K_i, K_j) = FDM(0, 0) minus the K_i x K_j block (which FDM never counts)
    rng = np.random.default_rng(17)
    di, dj, Ki, Kj = 13, 11, 5, 4
    base = random_pair(rng, di, dj, 0, 0)
    rho_i = np.diag(np.exp(-base.Ei[:Ki] / T) * WNF)
    rho_j = np.diag(np.exp(-base.Ej[:Kj] / T) * WNF)
    kd = Pair(base.Ei, base.Ej, base.op1[:, :di], base.op2[:, :di], Ki, Kj, rho_i, rho_j, base.factor)
    kk = Pair(base.Ei[:Ki], base.Ej[:Kj], base.op1[:Kj, :Ki], base.op2[:Kj, :Ki], 0, 0, factor=base.factor)
    g = grid(-4.0, 20.0, log10emax=1.0)
    full = run([base], g)
    part = run([kd, kk], g)
    scale = sum(np.abs(a).sum() for a in full)
    for a, b in zip(full, part):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14 * scale)


def test_total_weight_is_trace_form():
    # sum of all gt bins (no discard) = factor wnf sum_i exp(-Ei/T) (op1^T op2)_ii over all-discarded states
    rng = np.random.default_rng(3)
    p = random_pair(rng, 17, 23, 0, 0, emax=1.0)
    g = grid(-5.0, 10.0, log10emax=1.0)
    gtp, gtn, lsp, lsn = run([p], g, sign=-1)
    M = p.op1[:, :p.di].T @ p.op2[:, :p.di]
    expect_gt = p.factor * WNF * float(np.sum(np.exp(-p.Ei / T) * np.diag(M)))
    N = p.op1[:, :p.di] @ p.op2[:, :p.di].T
    expect_ls = p.factor * WNF * float(np.sum(np.exp(-p.Ej / T) * np.diag(N)))
    assert gtp.sum() + gtn.sum() == pytest.approx(expect_gt, rel=1e-12)
    assert lsp.sum() + lsn.sum() == pytest.approx(expect_ls, rel=1e-12)


def test_padding_of_operator_rows_is_not_read():
    rng = np.random.default_rng(9)
    a = random_pair(rng, 6, 5, 2, 3, pad=0)
    b = Pair(a.Ei, a.Ej, a.op1, a.op2, a.Ki, a.Kj, a.rho_i, a.rho_j, a.factor, pad=3)
    g = grid()
    for x, y in zip(run([a], g), run([b], g)):
        np.testing.assert_array_equal(x, y)
