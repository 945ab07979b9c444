"""Subspace pairs for the FDM binning tests (host arrays; device copies made by the GPU tests)."""
import numpy as np

from nrgljubljana_amd.core import BinGrid, FdmPair


def grid(log10emin=-6.0, bpd=10.0, nbins=None, log10emax=2.0, discard=0.0):
    """Bins::loggrid_std: nrbins = (log10emax - log10emin) * bins + 1 (c++/bins.hpp:95-100)."""
    if nbins is None:
        nbins = int((log10emax - log10emin) * bpd + 1.0)
    return BinGrid(log10emin, bpd, nbins, discard)


class Pair:
    """One (Ii, Ij) pair: op1, op2 are dim_j x dim_i (<j|op|i>), stored row-major with ld >= dim_i."""

    def __init__(self, Ei, Ej, op1, op2, Ki, Kj, rho_i=None, rho_j=None, factor=1.0, pad=0):
        self.Ei = np.ascontiguousarray(Ei, dtype=np.float64)
        self.Ej = np.ascontiguousarray(Ej, dtype=np.float64)
        dj, di = op1.shape
        self.ld = di + pad
        self.op1 = np.zeros((dj, self.ld))
        self.op2 = np.zeros((dj, self.ld))
        self.op1[:, :di] = op1
        self.op2[:, :di] = op2
        if pad:   # padding must never be read
            self.op1[:, di:] = np.nan
            self.op2[:, di:] = np.nan
        self.Ki, self.Kj = int(Ki), int(Kj)
        self.rho_i = None if rho_i is None else np.ascontiguousarray(rho_i, dtype=np.float64)
        self.rho_j = None if rho_j is None else np.ascontiguousarray(rho_j, dtype=np.float64)
        self.factor = float(factor)
        self.di, self.dj = di, dj

    def arrays(self):
        return [a for a in (self.Ei, self.Ej, self.op1, self.op2, self.rho_i, self.rho_j) if a is not None]

    def to_struct(self, ptr):
        """ptr(array) -> address; arrays may be numpy (oracle) or their device copies."""
        nri = 0 if self.rho_i is None else self.rho_i.shape[0]
        nrj = 0 if self.rho_j is None else self.rho_j.shape[0]
        return FdmPair(self.di, self.dj, self.Ki, self.Kj, ptr(self.Ei), ptr(self.Ej), ptr(self.op1), ptr(self.op2),
                       self.ld, ptr(self.rho_i) if self.rho_i is not None else None, nri,
                       ptr(self.rho_j) if self.rho_j is not None else None, nrj, self.factor)


def np_ptr(a):
    return a.ctypes.data


def random_pair(rng, di, dj, Ki, Kj, nri=None, nrj=None, emax=2.0, pad=0, factor=None):
    """Random pair with ascending energies in [0, emax) and random (non-symmetric) density matrices."""
    Ei = np.sort(rng.uniform(0, emax, di))
    Ej = np.sort(rng.uniform(0, emax, dj))
    op1 = rng.standard_normal((dj, di))
    op2 = rng.standard_normal((dj, di))
    nri = Ki if nri is None else nri
    nrj = Kj if nrj is None else nrj
    rho_i = rng.standard_normal((nri, nri)) / max(nri, 1) if nri else None
    rho_j = rng.standard_normal((nrj, nrj)) / max(nrj, 1) if nrj else None
    f = rng.uniform(0.5, 2.0) if factor is None else factor
    return Pair(Ei, Ej, op1, op2, Ki, Kj, rho_i, rho_j, f, pad)


def random_step(seed, npairs=24, dmax=300, emax=2.0):
    """A step's worth of pairs: ragged dims (incl. 0 and 1), K from 0 to dim, padded ld, larger rho."""
    rng = np.random.default_rng(seed)
    pairs = []
    for p in range(npairs):
        di = int(rng.integers(0, dmax + 1)) if p % 7 else int(rng.integers(0, 3))
        dj = int(rng.integers(1, dmax + 1))
        Ki = int(rng.integers(0, di + 1)) if p % 5 else 0
        Kj = int(rng.integers(0, dj + 1)) if p % 6 else dj
        nri = Ki + int(rng.integers(0, 3)) * (p % 3 == 0)
        nrj = Kj + int(rng.integers(0, 3)) * (p % 4 == 0)
        pairs.append(random_pair(rng, di, dj, Ki, Kj, nri, nrj, emax, pad=int(p % 3)))
    return pairs
