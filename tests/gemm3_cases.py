"""Term tables for the grouped triple products (shared by the CPU and GPU tests)."""
import ctypes

import numpy as np

from nrgljubljana_amd.core import Gemm3Term


class Mat:
    """A row-major matrix living inside a wider buffer (ld >= cols), like a ublas submatrix view."""

    def __init__(self, rng, rows, cols, pad=0, fill=True):
        self.rows, self.cols, self.ld = rows, cols, cols + pad
        self.buf = rng.standard_normal(rows * self.ld) if fill else np.zeros(rows * self.ld)
        self.dev = None

    def view(self):
        return self.buf.reshape(self.rows, self.ld)[:, :self.cols]

    def ptr(self, dev):
        if dev:
            return ctypes.c_void_p(self.dev.data_ptr())
        return self.buf.ctypes.data_as(ctypes.c_void_p)


def term(kind, C, A, B, O=None, f=1.0):
    """kind: product (M += f A B^T), transform (M += f A O B^T), rotate (M += f U^T O U)."""
    return dict(kind=kind, C=C, A=A, B=B, O=O, f=f)


def to_struct(t, dev):
    C, A, B, O = t["C"], t["A"], t["B"], t["O"]
    s = Gemm3Term()
    s.f = t["f"]
    s.m, s.n = C.rows, C.cols
    if t["kind"] == "product":
        s.k = s.l = A.cols
        s.transA, s.transB, s.O = 0, 0, None
    elif t["kind"] == "transform":
        s.k, s.l = O.rows, O.cols
        s.transA, s.transB = 0, 0
        s.O, s.ldo = O.ptr(dev), O.ld
    else:  # rotate: U = A = B is k x m
        s.k, s.l = O.rows, O.cols
        s.transA, s.transB = 1, 1
        s.O, s.ldo = O.ptr(dev), O.ld
    s.A, s.lda = A.ptr(dev), A.ld
    s.B, s.ldb = B.ptr(dev), B.ld
    s.C, s.ldc = C.ptr(dev), C.ld
    return s


def exact(terms, C):
    """Long-double evaluation of all terms that target C (the same sums, by definition)."""
    acc = C.view().astype(np.longdouble)
    for t in terms:
        if t["C"] is not C:
            continue
        A = t["A"].view().astype(np.longdouble)
        B = t["B"].view().astype(np.longdouble)
        f = np.longdouble(t["f"])
        if t["kind"] == "product":
            acc = acc + f * (A @ B.T)
        elif t["kind"] == "transform":
            acc = acc + f * (A @ t["O"].view().astype(np.longdouble) @ B.T)
        else:
            acc = acc + f * (A.T @ t["O"].view().astype(np.longdouble) @ B)
    return acc


def random_table(rng, nout=6, maxdim=60, pad=3):
    """Outputs C_g with several terms each, mixing the three kinds, with padded leading dimensions."""
    outs, terms = [], []
    for g in range(nout):
        m, n = int(rng.integers(1, maxdim)), int(rng.integers(1, maxdim))
        C = Mat(rng, m, n, pad)
        outs.append(C)
        for _ in range(int(rng.integers(1, 5))):
            kind = ["product", "transform", "rotate"][int(rng.integers(0, 3))]
            if kind == "rotate" and m != n:
                kind = "transform"
            f = float(rng.uniform(-2, 2))
            if kind == "product":
                k = int(rng.integers(1, maxdim))
                terms.append(term(kind, C, Mat(rng, m, k, pad), Mat(rng, n, k, pad), f=f))
            elif kind == "transform":
                k, l = int(rng.integers(1, maxdim)), int(rng.integers(1, maxdim))
                terms.append(term(kind, C, Mat(rng, m, k, pad), Mat(rng, n, l, pad), Mat(rng, k, l, pad), f=f))
            else:
                k = int(rng.integers(1, maxdim))
                U = Mat(rng, k, m, pad)
                terms.append(term(kind, C, U, U, Mat(rng, k, k, pad), f=f))
    return outs, terms
