"""Hamiltonian-assembly terms for the tests (host arrays; device copies made by the GPU tests)."""
import numpy as np

from nrgljubljana_amd.core import HamTerm, NRG_HAM_BLOCK, NRG_HAM_DIAG, NRG_HAM_IDENT


class Term:
    """One nrg_ham_term with its source array (None for IDENT and constant DIAG terms)."""

    def __init__(self, kind, block, row, col, rows, cols, src=None, factor=1.0, shift=0.0, trans=0, pad=0):
        self.kind, self.block, self.row, self.col, self.rows, self.cols = kind, block, row, col, rows, cols
        self.factor, self.shift, self.trans = float(factor), float(shift), int(trans)
        self.src = None
        self.ld = 0
        if src is not None:
            src = np.asarray(src, dtype=np.float64)
            if src.ndim == 1:
                self.src = np.ascontiguousarray(src)
            else:
                r, c = src.shape
                self.ld = c + pad
                self.src = np.full((r, self.ld), np.nan) if pad else np.zeros((r, self.ld))
                self.src[:, :c] = src

    def to_struct(self, ptr):
        return HamTerm(self.kind, self.trans, self.block, self.row, self.col, self.rows, self.cols,
                       ptr(self.src) if self.src is not None else None, self.ld, self.factor, self.shift)


def block_term(block, row, col, mat, factor, trans=False, pad=0):
    """offdiag_function: factor * mat at (row, col), or factor * mat^T when trans (then mat is cols x rows)."""
    mat = np.asarray(mat, dtype=np.float64)
    rows, cols = (mat.shape[1], mat.shape[0]) if trans else mat.shape
    return Term(NRG_HAM_BLOCK, block, row, col, rows, cols, mat, factor, trans=int(trans), pad=pad)


def diag_term(block, row, values=None, factor=1.0, shift=0.0, count=None):
    rows = len(values) if values is not None else count
    return Term(NRG_HAM_DIAG, block, row, row, rows, rows, values, factor, shift)


def ident_term(block, row, col, count, factor):
    return Term(NRG_HAM_IDENT, block, row, col, count, count, None, factor)


def arena(sizes):
    off = np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.int64) ** 2)[:-1]]).astype(np.int64)
    return off, int(sum(int(n) * int(n) for n in sizes))


def random_step(seed, nblk=12, nmax=300, nterms_per=10):
    """Blocks split into 2-6 sub-blocks (combs); the energies of every sub-block, zeta shifts, and
    offdiag blocks above the diagonal (both orientations of the source), identity couplings of
    equal-sized sub-blocks, sources padded (the padding must never be read)."""
    rng = np.random.default_rng(seed)
    sizes, terms = [], []
    for b in range(nblk):
        combs = int(rng.integers(2, 7))
        dims = [int(rng.integers(0, nmax // combs + 2)) for _ in range(combs)]
        if b % 5 == 0:
            dims = [d for d in dims] + [dims[0]]   # a repeated dimension for an identity coupling
        n = sum(dims)
        if n == 0:
            dims[0] = 1
            n = 1
        sizes.append(n)
        offs = np.concatenate([[0], np.cumsum(dims)[:-1]])
        for i, d in enumerate(dims):
            if d:
                terms.append(diag_term(b, int(offs[i]), rng.uniform(0, 10, d), factor=rng.uniform(0.5, 2)))
        for _ in range(nterms_per):
            i, j = sorted(rng.choice(len(dims), 2, replace=False))
            if dims[i] == 0 or dims[j] == 0:
                continue
            if rng.uniform() < 0.5:
                terms.append(block_term(b, int(offs[i]), int(offs[j]), rng.standard_normal((dims[i], dims[j])),
                                        rng.standard_normal(), pad=int(rng.integers(0, 3))))
            else:   # a term given for (j, i): its matrix is dims[j] x dims[i], added transposed
                terms.append(block_term(b, int(offs[i]), int(offs[j]), rng.standard_normal((dims[j], dims[i])),
                                        rng.standard_normal(), trans=True, pad=int(rng.integers(0, 3))))
        for i, d in enumerate(dims):
            if d and rng.uniform() < 0.5:
                terms.append(diag_term(b, int(offs[i]), None, shift=rng.standard_normal(), count=d))
        for i in range(len(dims)):
            for j in range(i + 1, len(dims)):
                if dims[i] and dims[i] == dims[j]:
                    terms.append(ident_term(b, int(offs[i]), int(offs[j]), dims[i], rng.standard_normal()))
    return sizes, terms
