"""Initial data of an NRG run: the `data` file (format version 9) of the single-impurity Anderson
model coupled to one or two channels, for the QS symmetry (charge U(1) x spin SU(2)).

NRG Ljubljana prepares this file with Mathematica (nrginit/initial.m, nrginit/models.m); this module
builds the same content from the model's definition:

* the impurity level d and the zeroth site f_c of every channel c, as a Fock space with
  Jordan-Wigner signs;
* H_0 = delta n_d + U/2 (n_d - 1)^2 + sum_c V_c sum_sigma (f_c,sigma^dag d_sigma + h.c.), the SIAM of
  nrginit/initial.m:509 with the hopping to the chain of nrginit/models.m:731-741, and
  V_c = sqrt(theta_c Gamma_c A / pi), theta_c = 2 for a flat band of half-width 1 (the integral of
  the hybridisation function over the band, nrginit/initial.m:465-466, 2898) and A = A_Lambda of the
  discretisation (nrginit/initial.m:996: (1 + 1/Lambda) ln(Lambda) / (2 (1 - 1/Lambda)) for Wilson's
  scheme "Y", 1 otherwise);
* the invariant subspaces (Q, SS): Q = N - (channels + 1) (charge from half filling), SS = 2S + 1,
  each represented by its highest-weight states S_z = S, with ascending energies
  (E - E_GS) / SCALE(0) and the absolute ground-state energy in the `e` block;
* reduced matrix elements under the Wigner-Eckart convention of the reference's recalculation tables:
  <j1 m1| T_q |j2 m2> = <j2 m2; k q | j1 m1> <j1||T||j2> (Condon-Shortley Clebsch-Gordan coefficients),
  for the f_c^dag of each channel (doublets, the `f` blocks), the doublet d^dag (`A_d`), the
  singlets n_d and n_d^2, and the triplet of the impurity spin S_d (`sigma_d`, its q = 0 component S_z).

The Wilson-chain coefficients come from the analytic flat-band formula of the discretisation "Y"
(Wilson, with twist z = 1):
    xi_n = (1 + 1/Lambda)(1 - Lambda^(-n-1)) / (2 sqrt((1 - Lambda^(-2n-1))(1 - Lambda^(-2n-3)))) Lambda^(-n/2),
    zeta_n = 0,
or are passed in (as the z block of an existing data file)."""
import math

import numpy as np


# ------------------------------------------------------------------ Fock space
class Fock:
    """Spin orbitals (site, spin) ordered as given; creation operators with Jordan-Wigner signs."""

    def __init__(self, nsites):
        self.n = 2 * nsites
        self.dim = 1 << self.n
        self.c = [self._creation(k) for k in range(self.n)]

    def _creation(self, k):
        m = np.zeros((self.dim, self.dim))
        for s in range(self.dim):
            if not (s >> k) & 1:
                sign = (-1) ** bin(s & ((1 << k) - 1)).count("1")
                m[s | (1 << k), s] = sign
        return m

    def cdag(self, site, spin):   # spin 0 = up, 1 = down
        return self.c[2 * site + spin]

    def number(self, site):
        a, b = self.cdag(site, 0), self.cdag(site, 1)
        return a @ a.T + b @ b.T


def _spin_ops(F, nsites):
    Sz = np.zeros((F.dim, F.dim))
    Sp = np.zeros((F.dim, F.dim))
    for s in range(nsites):
        u, d = F.cdag(s, 0), F.cdag(s, 1)
        Sz += 0.5 * (u @ u.T - d @ d.T)
        Sp += u @ d.T
    return Sz, Sp


def wilson_xi(Lam, nmax):
    """Flat-band Wilson-chain hoppings xi_0..xi_nmax (discretisation "Y", z = 1)."""
    n = np.arange(nmax + 1, dtype=float)
    return (1 + 1 / Lam) * (1 - Lam ** (-n - 1)) / (2 * np.sqrt((1 - Lam ** (-2 * n - 1)) * (1 - Lam ** (-2 * n - 3)))) \
        * Lam ** (-n / 2)


def scale(N, Lam, discretization="Y", z=1.0):
    """SCALE(N) of c++/params.hpp:619-634 (no substeps, bandrescale 1)."""
    s = 0.5 * (1 + 1 / Lam) if discretization == "Y" else (1 - 1 / Lam) / math.log(Lam)
    return s * Lam ** (-(N - 1) / 2 + 1 - z)


def _cg_half(j2, q, j1):
    """<j2 j2; 1/2 q | j1 j1> for j1 = j2 +- 1/2 and the q that reaches m1 = j1 from m2 = j2."""
    if abs(j1 - (j2 + 0.5)) < 1e-9:
        return 1.0                                    # q = +1/2
    return math.sqrt(2 * j2 / (2 * j2 + 1))           # j1 = j2 - 1/2, q = -1/2


def _cg_one_zero(j2, m, j1):
    """<j2 m; 1 0 | j1 m> (Condon-Shortley) for j1 in {j2 - 1, j2, j2 + 1}."""
    if abs(j1 - j2) < 1e-9:
        return m / math.sqrt(j2 * (j2 + 1)) if j2 > 0 else 0.0
    if abs(j1 - (j2 + 1)) < 1e-9:
        return math.sqrt((j2 - m + 1) * (j2 + m + 1) / ((2 * j2 + 1) * (j2 + 1)))
    return -math.sqrt((j2 - m) * (j2 + m) / (j2 * (2 * j2 + 1)))


def siam_qs(U, delta, gammas, Lam, discretization="Y", z=1.0, tol=1e-9):
    """Multiplets and reduced matrix elements of the initial cluster (impurity + site 0 of each channel).

    Returns dict(subspaces=[((Q, SS), energies)], GS=absolute GS energy, scale0=SCALE(0),
                 f=[{(I1, I2): matrix} per channel], ops={('s'|'d'|'t', name): {(I1, I2): matrix}})."""
    ch = len(gammas)
    nsites = 1 + ch                      # site 0 = d, sites 1..ch = f_c
    F = Fock(nsites)
    A = 0.5 * (1 + 1 / Lam) / (1 - 1 / Lam) * math.log(Lam) if discretization == "Y" else 1.0
    nd = F.number(0)
    I = np.eye(F.dim)
    H = delta * nd + U / 2 * (nd - I) @ (nd - I)
    for c, g in enumerate(gammas):
        V = math.sqrt(2.0 * g * A / math.pi)
        for sp in (0, 1):
            hop = F.cdag(1 + c, sp) @ F.cdag(0, sp).T
            H += V * (hop + hop.T)
    Ntot = sum(F.number(s) for s in range(nsites))
    Sz, Sp = _spin_ops(F, nsites)
    Nd, Szd = np.diag(Ntot), np.diag(Sz)
    # highest-weight states of every (Q, S): S_z = S and S^+ |psi> = 0, then H diagonalised there
    blocks = {}
    for N in range(F.dim.bit_length()):
        for twoSz in range(-nsites, nsites + 1):
            idx = np.where((np.abs(Nd - N) < 0.5) & (np.abs(2 * Szd - twoSz) < 0.5))[0]
            if not len(idx) or twoSz < 0:
                continue
            S = twoSz / 2
            # kernel of S^+ restricted to this sector
            Spb = Sp[:, idx]
            _, sv, vt = np.linalg.svd(Spb, full_matrices=True)
            null = vt[np.sum(sv > tol):].T if len(sv) else np.eye(len(idx))
            if null.shape[1] == 0:
                continue
            Hs = null.T @ H[np.ix_(idx, idx)] @ null
            e, v = np.linalg.eigh(0.5 * (Hs + Hs.T))
            vecs = np.zeros((F.dim, len(e)))
            vecs[idx, :] = null @ v
            Q = N - nsites
            blocks[(Q, int(round(2 * S)) + 1)] = (e, vecs)
    Egs = min(b[0][0] for b in blocks.values())
    s0 = scale(0, Lam, discretization, z)
    keys = sorted(blocks)

    def lower(vecs, j, m_from, steps):
        """(S^-)^steps |j m_from>, normalised as Condon-Shortley states."""
        v = vecs
        m = m_from
        for _ in range(steps):
            v = Sp.T @ v / math.sqrt(j * (j + 1) - m * (m - 1))
            m -= 1
        return v

    def doublet(opdag_up, opdag_dn):
        out = {}
        for (Q2, SS2), (e2, v2) in blocks.items():
            j2 = (SS2 - 1) / 2
            for SS1 in (SS2 + 1, SS2 - 1):
                k1 = (Q2 + 1, SS1)
                if k1 not in blocks or SS1 < 1:
                    continue
                j1 = (SS1 - 1) / 2
                op = opdag_up if SS1 == SS2 + 1 else opdag_dn
                m = blocks[k1][1].T @ op @ v2
                out[(k1, (Q2, SS2))] = m / _cg_half(j2, 0, j1)
        return out

    def singlet(op):
        return {(k, k): v.T @ op @ v for k, (e, v) in blocks.items()}

    def triplet(op_z):
        out = {}
        for (Q2, SS2), (e2, v2) in blocks.items():
            j2 = (SS2 - 1) / 2
            for SS1 in (SS2 - 2, SS2, SS2 + 2):
                k1 = (Q2, SS1)
                if k1 not in blocks or SS1 < 1:
                    continue
                j1 = (SS1 - 1) / 2
                m = min(j1, j2)   # a projection both multiplets have, with a nonzero <j2 m; 1 0 | j1 m>
                if abs(j1 - j2) < 1e-9 and m == 0:   # singlets: the element vanishes (written as zeros)
                    out[(k1, (Q2, SS2))] = np.zeros((blocks[k1][1].shape[1], v2.shape[1]))
                    continue
                b1 = lower(blocks[k1][1], j1, j1, int(round(j1 - m)))
                b2 = lower(v2, j2, j2, int(round(j2 - m)))
                out[(k1, (Q2, SS2))] = (b1.T @ op_z @ b2) / _cg_one_zero(j2, m, j1)
        return out

    f = [doublet(F.cdag(1 + c, 0), F.cdag(1 + c, 1)) for c in range(ch)]
    du, dd = F.cdag(0, 0), F.cdag(0, 1)
    ops = {("s", "n_d"): singlet(nd), ("s", "n_d^2"): singlet(nd @ nd), ("d", "A_d"): doublet(du, dd),
           ("t", "sigma_d"): triplet(0.5 * (du @ du.T - dd @ dd.T))}
    return dict(subspaces=[(k, (blocks[k][0] - Egs) / s0) for k in keys], GS=Egs, scale0=s0, f=f, ops=ops,
                channels=ch)


def _fmt(x):
    return repr(float(x))


def write_data(path, cluster, xi, zeta, ops=("n_d", "A_d")):
    """The `data` file (format 9): header, subspace energies, f blocks, e, operators, z block.
    xi, zeta: one table per channel, entries 0..Nmax."""
    ch = cluster["channels"]
    Nmax = len(xi[0]) - 1
    L = ["# Input file for NRG Ljubljana", "# symtype \tQS", "#!9",
         "# Number of channels, chain sites, subspaces: ",
         f"{ch}\t{Nmax}\t{len(cluster['subspaces'])}",
         f"# SCALE \t{_fmt(cluster['scale0'])}",
         "# Energies (GS energy subtracted, multiplied by 1/SCALE):"]
    dims = {}
    for (Q, SS), e in cluster["subspaces"]:
        dims[(Q, SS)] = len(e)
        L += [f"{Q}\t{SS}", str(len(e)), "\t".join(_fmt(x) for x in e)]

    def block(m):
        out = [str(len(m))]
        for (I1, I2), mat in sorted(m.items(), reverse=True):
            out.append(f"{I1[0]}\t{I1[1]}\t{I2[0]}\t{I2[1]}")
            out += ["\t".join(_fmt(x) for x in row) for row in np.atleast_2d(mat)]
        return out

    L.append("# Irreducible matrix elements for Wilson chains:")
    for c in range(ch):
        L.append(f"f {c} 0")
        L += block(cluster["f"][c])
    L += ["# GS energy in absolute units:", "e", _fmt(cluster["GS"]), "# Irreducible matrix elements for other operators:"]
    for (kind, name), m in cluster["ops"].items():
        if name in ops:
            L.append(f"{kind}{name}")
            L += block(m)
    L.append("# Discretization tables:")
    L.append("z")
    for t in list(xi) + list(zeta):
        L.append(str(len(t) - 1))
        L += [_fmt(x) for x in t]
    with open(path, "w") as F:
        F.write("\n".join(L) + "\n")


def read_coefficients(path, channels):
    """The xi and zeta tables (z block) of an existing data file: ([xi per channel], [zeta per channel])."""
    text = open(path).read()
    toks = text.split("\nz\n", 1)[1].split()
    vals = [float(t) for t in toks]
    tabs, i = [], 0
    while i < len(vals) and len(tabs) < 2 * channels:
        n = int(vals[i])
        tabs.append(np.array(vals[i + 1:i + 2 + n]))
        i += n + 2
    return tabs[:channels], tabs[channels:2 * channels]


def siam_data(path, U, delta, gammas, Lam, Nmax, ops=("n_d", "A_d")):
    """SIAM data file with flat-band Wilson coefficients (discretisation "Y")."""
    cl = siam_qs(U, delta, gammas, Lam, "Y")
    xi = [wilson_xi(Lam, Nmax)] * len(gammas)
    zeta = [np.zeros(Nmax + 1)] * len(gammas)
    write_data(path, cl, xi, zeta, ops)
    return cl
