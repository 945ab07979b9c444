"""Pins of the CPU oracle's eigensolver (oracle/nrg_oracle.c: dsyev as diagonalise_dsyev,
c++/diag.hpp:61-82) against the reference's known answers and against mathematics."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import compare_eig
from nrgljubljana_amd import workloads

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SQ2 = 1.0 / np.sqrt(2.0)


def test_sigma_z_known_answer():
    # test/unit/diag_test/diag_test.cpp:74-87
    w, V = oracle.syev(np.array([[-1.0, 0.0], [0.0, 1.0]]))
    assert w[0] == -1.0 and w[1] == 1.0
    assert abs(V[0, 0]) == 1.0 and V[0, 1] == 0.0 and V[1, 0] == 0.0 and abs(V[1, 1]) == 1.0


def test_sigma_x_known_answer():
    # test/unit/diag_test/diag_test.cpp:88-103 (EXPECT_DOUBLE_EQ = within 4 ulp)
    w, V = oracle.syev(np.array([[0.0, 1.0], [0.0, 0.0]]))
    assert w == pytest.approx([-1.0, 1.0], rel=1e-15, abs=1e-15)
    assert abs(V[0, 0]) == pytest.approx(SQ2, rel=1e-15) and abs(V[0, 1]) == pytest.approx(SQ2, rel=1e-15)
    assert V[0, 0] * V[0, 1] == pytest.approx(-0.5, rel=1e-15)
    assert V[1, 0] * V[1, 1] == pytest.approx(+0.5, rel=1e-15)


def test_diag1_golden_files():
    # test/tools/diag/diag1: mat.dat -> ref/val.dat, ref/vec.dat
    H = np.loadtxt(os.path.join(GOLD, "diag1", "mat.dat"))
    val = np.loadtxt(os.path.join(GOLD, "diag1", "val.dat"))
    vec = np.loadtxt(os.path.join(GOLD, "diag1", "vec.dat"))
    w, V = oracle.syev(np.triu(H))
    np.testing.assert_allclose(w, val, atol=1e-15)
    np.testing.assert_allclose(V, vec, atol=1e-15)


def test_lower_triangle_is_ignored():
    # the reference passes the row-major upper triangle as UPLO='L' (c++/diag.hpp:67); garbage below
    # the diagonal must not change anything
    rng = np.random.default_rng(3)
    h, _ = workloads.eig_block(rng, 30)
    g = h + np.tril(rng.standard_normal((30, 30)), -1)
    w1, V1 = oracle.syev(h)
    w2, V2 = oracle.syev(g)
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(V1, V2)


@pytest.mark.parametrize("n", [7, 50, 200])
def test_closed_form_spectrum(n):
    rng = np.random.default_rng(n)
    h, lam = workloads.eig_block(rng, n)
    w, V = oracle.syev(h)
    assert np.max(np.abs(w - lam)) <= 1e-13 * 20
    m = compare_eig(h, w, V, w, V)
    assert m["res"] < 1e-14 and m["orth"] < 1e-13


@pytest.mark.parametrize("n", [3, 12, 33])
def test_long_double_jacobi_brute_force(n):
    rng = np.random.default_rng(100 + n)
    h = np.triu(rng.standard_normal((n, n)))
    w, V = oracle.syev(h)
    wj, Vj = oracle.jacobi_ld(h)
    m = compare_eig(h, w, V, wj, Vj, hnorm=np.max(np.abs(wj)))
    assert m["eig"] < 1e-14 and m["proj"] < 1e-13


def test_batched_openmp_matches_serial():
    sizes, off_H, arena, off_w, lams = workloads.eig_batch(nblk=12, nmin=4, nmax=90, seed=5)
    a1, a2 = arena.copy(), arena.copy()
    w1, i1 = oracle.syev_batched(sizes, off_H, a1, off_w, threads=1)
    w2, i2 = oracle.syev_batched(sizes, off_H, a2, off_w, threads=4)
    assert not i1.any() and not i2.any()
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(a1, a2)
    for b, n in enumerate(sizes):
        assert np.max(np.abs(w1[off_w[b]:off_w[b] + n] - lams[b])) < 1e-12
