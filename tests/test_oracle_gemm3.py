"""Pins of the oracle's triple products (oracle/nrg_oracle.c gemm3: product / transform / rotate in
the reference's order, c++/numerics_ublas.hpp:108-140) against long-double brute force and invariants."""
import numpy as np
import pytest

import oracle
from tests.gemm3_cases import Mat, exact, random_table, term, to_struct


def run_oracle(terms):
    oracle.gemm3([to_struct(t, dev=False) for t in terms])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_tables_match_long_double(seed):
    rng = np.random.default_rng(seed)
    outs, terms = random_table(rng)
    ref = [exact(terms, C) for C in outs]
    run_oracle(terms)
    for C, R in zip(outs, ref):
        scale = float(np.max(np.abs(R))) + 1.0
        assert np.max(np.abs(C.view() - R.astype(np.float64))) <= 1e-13 * scale


def test_rotate_identity_gives_identity():
    # U orthogonal, O = I: U^T I U = I (recalculating the identity operator keeps it the identity)
    rng = np.random.default_rng(4)
    q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    U = Mat(rng, 30, 30)
    U.view()[:] = q
    O = Mat(rng, 30, 30)
    O.view()[:] = np.eye(30)
    C = Mat(rng, 30, 30, fill=False)
    run_oracle([term("rotate", C, U, U, O)])
    assert np.max(np.abs(C.view() - np.eye(30))) < 1e-14


def test_transposition_is_not_swapped():
    # A plausible mistake: A O B instead of A O B^T. With a non-square B the shapes differ, with a
    # square non-symmetric B the values differ; check the latter against hand-computed numbers.
    rng = np.random.default_rng(5)
    A, O, B = Mat(rng, 2, 2), Mat(rng, 2, 2), Mat(rng, 2, 2)
    A.view()[:] = [[1, 2], [3, 4]]
    O.view()[:] = [[0, 1], [1, 0]]
    B.view()[:] = [[1, 0], [5, 1]]
    C = Mat(rng, 2, 2, fill=False)
    run_oracle([term("transform", C, A, B, O, f=2.0)])
    # A O B^T = [[2,1],[4,3]] @ [[1,5],[0,1]] = [[2,11],[4,23]]
    np.testing.assert_array_equal(C.view(), 2.0 * np.array([[2, 11], [4, 23]]))


def test_empty_terms_contribute_nothing():
    rng = np.random.default_rng(6)
    C = Mat(rng, 4, 5)
    before = C.view().copy()
    run_oracle([term("product", C, Mat(rng, 4, 0), Mat(rng, 5, 0))])
    np.testing.assert_array_equal(C.view(), before)
