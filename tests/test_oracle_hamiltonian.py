"""Pins of the oracle's Hamiltonian assembly (hamiltonian(), c++/core.hpp:72-88; offdiag_function_impl,
diag_function_impl, diag_offdiag_function, c++/matrix.hpp:32-111).

The checks do not restate the loops: they compare spectra of the assembled matrices with closed
forms and with singular values of the coupling blocks, which fail for a block placed at the wrong
offset, a missing transpose, a dropped factor or a lost shift."""
import numpy as np
import pytest

import oracle
from tests.fdm_cases import np_ptr
from tests.ham_cases import arena, block_term, diag_term, ident_term


def assemble(sizes, terms):
    off, tot = arena(sizes)
    H = np.full(tot, np.nan)
    oracle.hamiltonian(sizes, off, H, [t.to_struct(np_ptr) for t in terms])
    return [H[o:o + n * n].reshape(n, n) for o, n in zip(off, sizes)]


def spectrum(h):
    """Eigenvalues of the symmetric matrix whose upper triangle h holds."""
    return np.linalg.eigvalsh(np.triu(h) + np.triu(h, 1).T)


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.lib()


def test_two_level_closed_form():
    d1, d2, t, scale = 0.3, 1.7, -0.45, 2.0
    (h,) = assemble([2], [diag_term(0, 0, [d1, d2], factor=scale), block_term(0, 0, 1, [[t]], 1.0)])
    assert np.all(np.tril(h, -1) == 0)
    a, b = scale * d1, scale * d2
    r = np.sqrt(((a - b) / 2) ** 2 + t * t)
    assert spectrum(h) == pytest.approx([(a + b) / 2 - r, (a + b) / 2 + r], rel=1e-15)


@pytest.mark.parametrize("m", [1, 7, 70])
def test_identity_coupling_closed_form(m):
    # diag_offdiag_function couples two sub-blocks of equal size: spectrum of [[a, t], [t, b]] m times
    a, b, t = 0.25, -0.5, 0.8
    terms = [diag_term(0, 0, None, shift=a, count=m), diag_term(0, m, None, shift=b, count=m),
             ident_term(0, 0, m, m, t)]
    (h,) = assemble([2 * m], terms)
    r = np.sqrt(((a - b) / 2) ** 2 + t * t)
    want = np.array([(a + b) / 2 - r] * m + [(a + b) / 2 + r] * m)
    assert spectrum(h) == pytest.approx(want, abs=1e-14)


@pytest.mark.parametrize("dims", [(3, 5, 4), (40, 33, 70), (1, 1, 1)])
def test_coupling_blocks_and_transpose(dims):
    # sub-blocks 1-2 coupled by A (term (1,2)) and 2-3 by a term given for (3,2), i.e. a matrix M of
    # shape r3 x r2 added as M^T into the upper block (2,3). The matrix is bipartite between {1,3}
    # and {2} with coupling X = [A; M]: its eigenvalues are +-sqrt(eig(X^T X)) = +-sqrt(eig(A^T A +
    # M^T M)) and zeros. Without the transpose X = [A; M^T] and the spectrum changes.
    r1, r2, r3 = dims
    rng = np.random.default_rng(sum(dims))
    A = rng.standard_normal((r1, r2))
    M = rng.standard_normal((r3, r2))
    fa, fm = 0.7, -1.3
    terms = [block_term(0, 0, r1, A, fa, pad=2), block_term(0, r1, r1 + r2, M, fm, trans=True, pad=1)]
    (h,) = assemble([r1 + r2 + r3], terms)
    s = np.sqrt(np.clip(np.linalg.eigvalsh((fa * A).T @ (fa * A) + (fm * M).T @ (fm * M)), 0, None))
    want = np.sort(np.concatenate([s, -s, np.zeros(r1 + r3 - r2)])) if r1 + r3 >= r2 else None
    got = spectrum(h)
    assert got == pytest.approx(want, abs=1e-12 * max(1.0, s.max()))


def test_shift_moves_the_whole_spectrum_and_energies_scale():
    rng = np.random.default_rng(3)
    dims = [6, 9]
    A = rng.standard_normal((6, 9))
    e1, e2 = rng.uniform(0, 5, 6), rng.uniform(0, 5, 9)
    base = [diag_term(0, 0, e1, factor=1.5), diag_term(0, 6, e2, factor=1.5), block_term(0, 0, 6, A, 0.4)]
    (h0,) = assemble([15], base)
    (h1,) = assemble([15], base + [diag_term(0, 0, None, shift=0.125, count=6),
                                   diag_term(0, 6, None, shift=0.125, count=9)])
    assert spectrum(h1) == pytest.approx(spectrum(h0) + 0.125, abs=1e-13)
    # the energies are scaled: with no coupling the spectrum is 1.5 * (e1, e2)
    (h2,) = assemble([15], base[:2])
    assert spectrum(h2) == pytest.approx(np.sort(1.5 * np.concatenate([e1, e2])), rel=1e-15)


def test_terms_on_the_same_elements_add_and_blocks_are_independent():
    rng = np.random.default_rng(4)
    A, B = rng.standard_normal((3, 4)), rng.standard_normal((3, 4))
    (h,) = assemble([7], [block_term(0, 0, 3, A, 1.0), block_term(0, 0, 3, B, 1.0)])
    (hs,) = assemble([7], [block_term(0, 0, 3, A + B, 1.0)])
    assert spectrum(h) == pytest.approx(spectrum(hs), abs=1e-13)
    # a term on block 1 does not touch block 0, and every element of every block is written
    h0, h1 = assemble([3, 4], [diag_term(1, 0, [1.0, 2.0, 3.0, 4.0])])
    assert np.all(h0 == 0) and np.all(np.isfinite(h1))
    assert spectrum(h1) == pytest.approx([1.0, 2.0, 3.0, 4.0], rel=1e-15)
