"""Seeded synthetic inputs of the per-step NRG core (host side, numpy).

eig_batch(): the batched-eigensolver workload -- real symmetric blocks H = Q diag(lambda) Q^T with
Haar-random Q (QR of a Gaussian matrix with the signs of diag(R) fixed), sizes log-uniform in
[16, 1024], eigenvalues sorted uniform in [0, 20] with, per block, a cluster of four values spaced
1e-10 and one exactly degenerate pair. Blocks are packed into one arena, row-major with the upper
triangle filled, as diagonalise() receives them (c++/diag.hpp:265).
"""
import numpy as np

SEED = 20261203


def eig_sizes(nblk=256, nmin=16, nmax=1024, seed=SEED):
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(nmin), np.log(nmax), size=nblk)
    return np.clip(np.rint(np.exp(u)), nmin, nmax).astype(np.int64)


def haar_orthogonal(rng, n):
    g = rng.standard_normal((n, n))
    q, r = np.linalg.qr(g)
    return q * np.sign(np.diag(r))[None, :]


def spectrum(rng, n, lmax=20.0):
    lam = np.sort(rng.uniform(0.0, lmax, size=n))
    if n >= 8:
        i = int(rng.integers(0, n - 6))
        lam[i:i + 4] = lam[i] + np.arange(4) * 1e-10     # cluster spaced 1e-10
        lam[i + 5] = lam[i + 4]                          # exactly degenerate pair
        lam = np.sort(lam)
    return lam


def eig_block(rng, n, lmax=20.0):
    q = haar_orthogonal(rng, n)
    lam = spectrum(rng, n, lmax)
    h = (q * lam[None, :]) @ q.T
    h = 0.5 * (h + h.T)
    return np.triu(h), lam


def eig_batch(nblk=256, nmin=16, nmax=1024, seed=SEED, sizes=None):
    """Returns (sizes, off_H, arena, off_w, lambdas): arena is float64 of length sum(n^2)."""
    if sizes is None:
        sizes = eig_sizes(nblk, nmin, nmax, seed)
    sizes = np.asarray(sizes, dtype=np.int64)
    rng = np.random.default_rng(seed + 1)
    off_H = np.zeros(len(sizes), np.int64)
    off_w = np.zeros(len(sizes), np.int64)
    if len(sizes) > 1:
        off_H[1:] = np.cumsum(sizes[:-1] ** 2)
        off_w[1:] = np.cumsum(sizes[:-1])
    arena = np.zeros(int(np.sum(sizes ** 2)), np.float64)
    lams = []
    for b, n in enumerate(sizes):
        h, lam = eig_block(rng, int(n))
        arena[off_H[b]:off_H[b] + n * n] = h.ravel()
        lams.append(lam)
    return sizes, off_H, arena, off_w, lams


def eig_flops(sizes):
    """Algorithmic flops of the eigensolve: (10/3) n^3 per block (reduction + back-transformation)."""
    s = np.asarray(sizes, dtype=np.float64)
    return float(np.sum(10.0 / 3.0 * s ** 3))


def eig_bytes(sizes):
    """Algorithmic HBM bytes: read the upper triangle, write the eigenvectors and eigenvalues."""
    s = np.asarray(sizes, dtype=np.float64)
    return float(np.sum(8.0 * (s * (s + 1) / 2 + s * s + s)))
