"""Seeded synthetic inputs of the per-step NRG core (host side, numpy).

eig_batch(): the batched-eigensolver workload -- real symmetric blocks H = Q diag(lambda) Q^T with
Haar-random Q (QR of a Gaussian matrix with the signs of diag(R) fixed), sizes log-uniform in
[16, 1024], eigenvalues sorted uniform in [0, 20] with, per block, a cluster of four values spaced
1e-10 and one exactly degenerate pair. Blocks are packed into one arena, row-major with the upper
triangle filled, as diagonalise() receives them (c++/diag.hpp:265).
"""
import os

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


def recalc_table(npairs=48, dmin=16, dmax=2000, combs=16, seed=SEED + 2):
    """The split-precision recalc workload: npairs output blocks <I1||O||Ip> with dims log-uniform in
    [dmin, dmax]; each subspace dimension is a random composition of `combs` ancestor dimensions (16
    for two-channel QSZ) and every output gets one table row per ancestor (the 16-row
    qsz-2ch-doublet tables: i1 = ip = i), i.e. the recalc_general term
        C(I1,Ip) += f_i U_I1(:, part_i) O_i U_Ip(:, part'_i)^T,  O_i ~ N(0,1) of r_i x r'_i,
    with U_I1, U_Ip Haar orthogonal and f_i = +-1. Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(dmin), np.log(dmax), size=(npairs, 2))
    dims = np.clip(np.rint(np.exp(u)), dmin, dmax).astype(np.int64)

    def composition(d):
        cuts = np.sort(rng.choice(np.arange(1, d), size=min(combs - 1, d - 1), replace=False)) if d > 1 else []
        parts = np.diff(np.concatenate([[0], cuts, [d]])).astype(np.int64)
        return np.concatenate([parts, np.zeros(combs - len(parts), np.int64)])

    out = []
    for dim1, dimp in dims:
        r1, rp = composition(int(dim1)), composition(int(dimp))
        U1, Up = haar_orthogonal(rng, int(dim1)), haar_orthogonal(rng, int(dimp))
        Os = [rng.standard_normal((int(a), int(b))) for a, b in zip(r1, rp)]
        f = rng.choice([-1.0, 1.0], size=combs)
        out.append(dict(dim1=int(dim1), dimp=int(dimp), r1=r1, rp=rp, U1=U1, Up=Up, O=Os, f=f))
    return out


def recalc_flops(table):
    """Algorithmic FP64 flops in the reference's order (c++/numerics_ublas.hpp:118-128):
    T = O U_p^T (2 r rp dimp), then C += U1 T (2 dim1 r dimp), per table row."""
    s1 = s2 = 0.0
    for t in table:
        for a, b in zip(t["r1"], t["rp"]):
            if a and b:
                s1 += 2.0 * a * b * t["dimp"]
                s2 += 2.0 * t["dim1"] * a * t["dimp"]
    return s1, s2


def fdm_step(npairs=96, dmin=16, dmax=2000, keep_frac=0.4, emax=10.0, seed=SEED + 3):
    """One FDM step's subspace pairs: dims log-uniform in [dmin, dmax], the lowest keep_frac of each
    subspace kept, ascending energies in [0, emax), op1 = op2 Gaussian (an A-A spectral function),
    density matrices random symmetric over the kept states."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(dmin), np.log(dmax), size=(npairs, 2))
    dims = np.clip(np.rint(np.exp(u)), dmin, dmax).astype(np.int64)
    pairs = []
    for di, dj in dims:
        Ki, Kj = int(round(keep_frac * di)), int(round(keep_frac * dj))
        Ei, Ej = np.sort(rng.uniform(0, emax, int(di))), np.sort(rng.uniform(0, emax, int(dj)))
        op = rng.standard_normal((int(dj), int(di)))
        ri = rng.standard_normal((Ki, Ki))
        rj = rng.standard_normal((Kj, Kj))
        pairs.append(dict(dim_i=int(di), dim_j=int(dj), K_i=Ki, K_j=Kj, E_i=Ei, E_j=Ej, op=op,
                          rho_i=(ri + ri.T) / max(Ki, 1), rho_j=(rj + rj.T) / max(Kj, 1)))
    return pairs


def fdm_elements(pairs):
    """Elements evaluated per half (gt or ls): |D_i||D_j| + |D_i||K_j| + |K_i||D_j|."""
    return sum(p["dim_i"] * p["dim_j"] - p["K_i"] * p["K_j"] for p in pairs)


def cfs_elements(pairs):
    """Elements a CFS step below the last one evaluates: |K_i||D_j| (Eq. 16) + |D_i||K_j| (Eq. 15)."""
    return sum(p["K_i"] * (p["dim_j"] - p["K_j"]) + p["K_j"] * (p["dim_i"] - p["K_i"]) for p in pairs)


def cfs_flops(pairs):
    """The two density-matrix products of a CFS step: op2[D_j, :] rho_i (2 |D_j| K_i^2) and
    rho_j^T op1[:, D_i] (2 |D_i| K_j^2), density matrices over the kept states."""
    return sum(2.0 * (p["dim_j"] - p["K_j"]) * p["K_i"] ** 2 + 2.0 * (p["dim_i"] - p["K_i"]) * p["K_j"] ** 2
               for p in pairs)


def ham_step(nblk=256, seed=SEED + 6):
    """One step's Hamiltonian blocks with the eigensolver workload's orders (log-uniform [16,1024]):
    each block split into 4 sub-blocks (the combs of a one-channel QS step), the ancestors' energies
    on the diagonal, <||f||> couplings between sub-blocks (1,2), (1,3), (2,4), (3,4) as in a
    one-channel table (the pair (3,4) given as (4,3), i.e. transposed), and a zeta shift per
    sub-block. Returns (sizes, blocks), blocks[b] = dict(dims, corr, mats)."""
    rng = np.random.default_rng(seed)
    sizes = eig_sizes(nblk, seed=SEED)
    blocks = []
    for n in sizes:
        n = int(n)
        cuts = np.sort(rng.choice(np.arange(1, n), 3, replace=False)) if n > 3 else np.array([1, 2, 3])[:max(n - 1, 0)]
        dims = np.diff(np.concatenate([[0], cuts, [n]])).astype(int).tolist()
        while len(dims) < 4:
            dims.append(0)
        corr = [np.sort(rng.uniform(0, 10, d)) for d in dims]
        mats = {}
        for (i, j) in ((1, 2), (1, 3), (2, 4), (4, 3)):
            if dims[i - 1] and dims[j - 1]:
                mats[(i, j)] = rng.standard_normal((dims[i - 1], dims[j - 1]))
        blocks.append(dict(dims=dims, corr=corr, mats=mats))
    return sizes, blocks


def dmnrg_step(npairs=48, kmin=16, kmax=1000, emax=10.0, seed=SEED + 5):
    """One DM-NRG spectral step's subspace pairs (Algo_DMNRG, kept x kept elements only): kept dims
    log-uniform in [kmin, kmax] (no discarded states), ascending abs_zero energies in [0, emax),
    op1 = op2 Gaussian, density matrices random symmetric positive with unit trace."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(kmin), np.log(kmax), size=(npairs, 2))
    dims = np.clip(np.rint(np.exp(u)), kmin, kmax).astype(np.int64)
    pairs = []
    for ki, kj in dims:
        ki, kj = int(ki), int(kj)
        Ei, Ej = np.sort(rng.uniform(0, emax, ki)), np.sort(rng.uniform(0, emax, kj))
        op = rng.standard_normal((kj, ki))
        rhos = []
        for k in (ki, kj):
            a = rng.standard_normal((k, k))
            m = a @ a.T
            rhos.append(m / np.trace(m))
        pairs.append(dict(dim_i=ki, dim_j=kj, K_i=ki, K_j=kj, E_i=Ei, E_j=Ej, op=op, rho_i=rhos[0], rho_j=rhos[1]))
    return pairs


def dmnrg_flops(pairs):
    """The two density-matrix products of every pair: 2 K_i K_j (K_i + K_j)."""
    return sum(2.0 * p["K_i"] * p["K_j"] * (p["K_i"] + p["K_j"]) for p in pairs)


def dm_step(nsub=48, combs=16, kmin=16, kmax=1000, seed=SEED + 4):
    """One backward step of the density-matrix sweep (cdmI, c++/dmnrg.hpp:54-80): nsub subspaces of
    step N-1 with kept dimensions log-uniform in [kmin, kmax]; each has `combs` children at step N
    whose ancestor block i has rmax = the parent's kept dimension; child c holds rho_N (nromega x
    nromega, nromega = its kept states) and eigenvectors U (nromega x its total dimension), of which
    the columns part(i) enter: rho_{N-1}(I) += f U(:, part)^T rho_N(child) U(:, part)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(kmin), np.log(kmax), size=nsub)
    kept = np.clip(np.rint(np.exp(u)), kmin, kmax).astype(np.int64)
    out = []
    for k in kept:
        children = []
        for c in range(combs):
            nrom = int(np.clip(np.rint(np.exp(rng.uniform(np.log(kmin), np.log(kmax)))), kmin, kmax))
            total = int(k) + int(rng.integers(0, 2 * kmax))   # the parent's block plus the other ancestors'
            off = int(rng.integers(0, total - int(k) + 1))
            children.append(dict(nromega=nrom, total=total, off=off, f=float(rng.uniform(0.3, 1.0))))
        out.append(dict(kept=int(k), children=children))
    return out


def dm_flops(step):
    """rotate in the reference's order (c++/numerics_ublas.hpp:130-140): T = U^T rho (2 rmax nromega^2),
    then M += f T U (2 rmax^2 nromega), per (subspace, child)."""
    s = 0.0
    for sub in step:
        r = sub["kept"]
        for ch in sub["children"]:
            s += 2.0 * r * ch["nromega"] ** 2 + 2.0 * r * r * ch["nromega"]
    return s


# ---- end-to-end NRG runs at the configurations' parameter points
NRG_CONFIGS = {
    # single-channel Anderson model, QS, Lambda = 2, U = 0.1, eps_d = -U/2, Gamma = 0.01, T = 1e-8,
    # keep <= 1000 multiplets with cut-off energy 10, DM-NRG spectral function of d
    "c1": dict(channels=1, U=0.1, Gamma=0.01, Lambda=2.0, T=1e-8, keep=1000, keepenergy=10.0,
               algo="dmnrg=true", ops="n_d A_d"),
    # two-channel Anderson model, QS, Lambda = 3, U = 0.2, Gamma_1 = Gamma_2 = 0.02, T = 1e-10, keep 4000
    # multiplets, FDM spectral function of d and FDM thermodynamics
    "c5": dict(channels=2, U=0.2, Gamma=0.02, Lambda=3.0, T=1e-10, keep=4000, keepenergy=-1.0,
               algo="fdm=true", ops="n_d A_d"),
}


def nrg_chain_length(Lam, T):
    """Chain length of a run down to temperature T: the first N with SCALE(N) <= T (discretisation Y)."""
    from .initial import scale
    N = 1
    while scale(N, Lam, "Y") > T:
        N += 1
    return N


def nrg_config(name, directory, diag="default", keep=None, nmax=None):
    """Writes `param` and `data` of configuration `name` (NRG_CONFIGS) into `directory`: flat-band
    Wilson chain (discretisation Y), eps_d = -U/2 (delta = 0), the initial cluster from
    initial.siam_qs. `keep` and `nmax` override the configuration (shorter runs for tests).
    Returns the chain length."""
    from . import initial
    c = NRG_CONFIGS[name]
    N = nmax or nrg_chain_length(c["Lambda"], c["T"])
    k = keep or c["keep"]
    os.makedirs(directory, exist_ok=True)
    cl = initial.siam_qs(c["U"], 0.0, [c["Gamma"]] * c["channels"], c["Lambda"], "Y")
    xi = [initial.wilson_xi(c["Lambda"], N)] * c["channels"]
    zeta = [np.zeros(N + 1)] * c["channels"]
    initial.write_data(os.path.join(directory, "data"), cl, xi, zeta, ops=tuple(c["ops"].split()))
    with open(os.path.join(directory, "param"), "w") as F:
        F.write("[param]\nsymtype=QS\ndiscretization=Y\n"
                f"Lambda={c['Lambda']}\nT={c['T']}\nkeep={k}\nkeepenergy={c['keepenergy']}\n"
                f"ops={c['ops']}\nspecd=A_d-A_d\n{c['algo']}\ndiag={diag}\n")
    return N
