"""Pins of the oracle's spectral broadening (SpectrumRealFreq::continuous, c++/spectrum.hpp:180-209;
BR_L / BR_G / BR_NEW, c++/spectral.hpp:39-71) by properties of the kernels, not by restating them:
normalisation, the argument symmetry that gamma = alpha / 4 gives, continuity at |E| = omega0,
conservation of a peak's weight, sign separation and linearity."""
import math

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def log_mesh(lo, hi, n):
    return np.exp(np.linspace(math.log(lo), math.log(hi), n))


@pytest.mark.parametrize("alpha", [0.3, 0.6, 0.1])
@pytest.mark.parametrize("ept", [1e-6, 0.2, 3.0])
def test_log_gaussian_is_normalised(alpha, ept):
    # int_0^inf BR_L(e, ept) de = 1: with x = ln(e/ept) the integrand is a unit Gaussian in x
    e = log_mesh(ept * math.exp(-12 * alpha), ept * math.exp(12 * alpha), 20001)
    f = np.array([oracle.br_l(x, ept, alpha) for x in e])
    assert np.trapezoid(f, e) == pytest.approx(1.0, abs=1e-7)
    # negative peaks mirror positive ones, and nothing crosses the sign boundary
    assert oracle.br_l(-e[100], -ept, alpha) == pytest.approx(f[100], rel=1e-15)
    assert oracle.br_l(-e[100], ept, alpha) == 0.0 and oracle.br_l(e[100], 0.0, alpha) == 0.0


def test_gaussian_is_normalised():
    w0 = 0.01
    e = np.linspace(-0.2, 0.2, 40001) + 0.003
    f = np.array([oracle.br_g(x, 0.003, w0) for x in e])
    assert np.trapezoid(f, e) == pytest.approx(1.0, abs=1e-9)


def test_log_gaussian_symmetric_in_its_arguments():
    # gamma = alpha/4 makes BR_L(e, e') = BR_L(e', e): (a+b)^2 - (a-b)^2 = 4ab = ln(e/e') cancels 1/e vs 1/e'
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b = 10 ** rng.uniform(-4, 1, 2)
        alpha = rng.uniform(0.1, 0.8)
        assert oracle.br_l(a, b, alpha) == pytest.approx(oracle.br_l(b, a, alpha), rel=1e-13)


def test_blend_is_continuous_and_log_gaussian_above_omega0():
    alpha, w0 = 0.3, 1e-3
    for ept in (2e-4, 1e-3, 5e-3):
        assert oracle.br_new(2.5 * w0, ept, alpha, w0) == oracle.br_l(2.5 * w0, ept, alpha)
        below = oracle.br_new(w0 * (1 - 1e-9), ept, alpha, w0)
        above = oracle.br_new(w0 * (1 + 1e-9), ept, alpha, w0)
        assert below == pytest.approx(above, rel=1e-6)


def test_single_peak_keeps_its_weight_and_sign():
    alpha, w0 = 0.3, 1e-8
    mesh = log_mesh(1e-4, 1e2, 6000)
    ap, an = oracle.broaden([0.5], [2.5], [], [], mesh, alpha, w0)
    assert np.trapezoid(ap, mesh) == pytest.approx(2.5, rel=1e-6)
    assert not an.any()                         # a positive peak never shows at negative energies
    bp, bn = oracle.broaden([], [], [0.5], [1.5], mesh, alpha, w0)
    assert np.trapezoid(bn, mesh) == pytest.approx(1.5, rel=1e-6) and not bp.any()


def test_low_energy_gaussian_crosses_zero():
    # below omega0 the Gaussian part reaches the other sign (BR_G has no sign condition)
    alpha, w0 = 0.3, 1e-2
    mesh = np.array([1e-3, 5e-3, 2e-2])
    ap, an = oracle.broaden([2e-3], [1.0], [], [], mesh, alpha, w0)
    assert an[0] > 0 and an[1] > 0 and an[2] == 0.0


def test_linearity():
    rng = np.random.default_rng(4)
    e1, w1 = 10 ** rng.uniform(-5, 1, 50), rng.standard_normal(50)
    e2, w2 = 10 ** rng.uniform(-5, 1, 40), rng.standard_normal(40)
    mesh = log_mesh(1e-5, 10, 300)
    a = oracle.broaden(e1, w1, e2, w2, mesh, 0.4, 1e-4)
    b1 = oracle.broaden(e1, w1, [], [], mesh, 0.4, 1e-4)
    b2 = oracle.broaden([], [], e2, w2, mesh, 0.4, 1e-4)
    for x, y, z in zip(a, b1, b2):
        np.testing.assert_allclose(x, y + z, rtol=1e-12, atol=1e-12 * np.abs(x).max())
