"""The initial-data generator (nrgljubljana_amd/initial.py) against the reference's own data files.

The reference's data files hold the multiplet energies, the ground-state energy and the reduced
matrix elements of the initial cluster, written by its Mathematica front end. The generator must
give the same energies and ground-state energy, and the same matrix elements up to the sign of each
multiplet state (the eigenvectors' gauge; the end-to-end GPU test pins the signs' consistency by
reproducing the fixtures' ref/ outputs from generated data). The analytic flat-band Wilson
coefficients are pinned by a Lanczos tridiagonalisation of the logarithmically discretised band."""
import os

import numpy as np
import pytest

from nrgljubljana_amd import initial as ini

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "nrg")

# fixture, channels, operators to check: U, Gamma, delta and Lambda come from its param file
CASES = [("c++__test0_minimal", 1), ("c++__test1_qs", 1), ("c++__legacy_substeps_one_siam_std", 2)]


def params(path):
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            kv[k] = v
    return kv


def read_data(path, channels):
    """(subspace energies, GS energy, named blocks) of a reference data file."""
    lines = [l.strip() for l in open(path)]
    body = [l for l in lines if l and not l.startswith("#")]
    toks = " ".join(body).split()
    pos = 0

    def take(n=1):
        nonlocal pos
        out = toks[pos:pos + n]
        pos += n
        return out

    ch, nmax, nsubs = (int(x) for x in take(3))
    energies = {}
    for _ in range(nsubs):
        q, ss = (int(x) for x in take(2))
        n = int(take()[0])
        energies[(q, ss)] = np.array([float(x) for x in take(n)])

    def block():
        nb = int(take()[0])
        out = {}
        for _ in range(nb):
            a, b, c, d = (int(x) for x in take(4))
            r, cc = len(energies[(a, b)]), len(energies[(c, d)])
            out[((a, b), (c, d))] = np.array([float(x) for x in take(r * cc)]).reshape(r, cc)
        return out

    blocks = {}
    for c in range(channels):
        assert take(3)[0] == "f"
        blocks[("f", c)] = block()
    GS = None
    while pos < len(toks):
        t = take()[0]
        if t == "e":
            GS = float(take()[0])
        elif t[0] in "sdt" and len(t) > 1:
            blocks[(t[0], t[1:])] = block()
        else:
            break
    return energies, GS, blocks


@pytest.mark.parametrize("name,channels", CASES)
def test_generator_reproduces_reference_data(name, channels):
    p = params(os.path.join(GOLDEN, name, "param"))
    gam = [float(p["Gamma"])] * channels
    cl = ini.siam_qs(float(p["U"]), float(p.get("delta", 0)), gam, float(p["Lambda"]), p.get("discretization", "Z")[0])
    energies, GS, blocks = read_data(os.path.join(GOLDEN, name, "data"), channels)
    mine = dict(cl["subspaces"])
    assert set(mine) == set(energies)
    for k, e in energies.items():
        assert np.allclose(mine[k], e, rtol=0, atol=1e-12), (k, mine[k], e)
    assert abs(cl["GS"] - GS) <= 1e-12
    ours = {("f", c): cl["f"][c] for c in range(channels)}
    ours.update({(kind, nm): m for (kind, nm), m in cl["ops"].items()})
    checked = 0
    for key, ref in blocks.items():
        assert key in ours, key
        for II, m in ref.items():
            assert II in ours[key], (key, II)
            assert np.allclose(np.abs(ours[key][II]), np.abs(m), rtol=0, atol=1e-12), (key, II)
            checked += 1
    assert checked > 10


def test_wilson_coefficients_from_lanczos():
    """Lanczos on the star form of Wilson's discretisation (intervals [L^-(m+1), L^-m] on both sides,
    energies (1 + 1/L) L^-m / 2, weights (1 - 1/L) L^-m / 2, normalised) gives the analytic xi_n."""
    for L in (2.0, 3.0):
        M = 80
        m = np.arange(M)
        eps = 0.5 * (1 + 1 / L) * L ** (-m)
        w = 0.5 * (1 - 1 / L) * L ** (-m)
        E = np.concatenate([eps, -eps])
        g = np.sqrt(np.concatenate([w, w]))
        v = g / np.linalg.norm(g)
        vprev, b, xi = np.zeros_like(v), 0.0, []
        for n in range(8):
            u = E * v - b * vprev
            a = v @ u
            u -= a * v
            b = np.linalg.norm(u)
            xi.append(b)
            vprev, v = v, u / b
        assert np.allclose(xi, ini.wilson_xi(L, 7), rtol=1e-10, atol=0), (L, xi)


def test_written_data_file_reads_back(tmp_path):
    cl = ini.siam_qs(0.2, 0.0, [0.02, 0.02], 3.0, "Y")
    xi, ze = [ini.wilson_xi(3.0, 10)] * 2, [np.zeros(11)] * 2
    ini.write_data(tmp_path / "data", cl, xi, ze, ops=("n_d", "A_d"))
    energies, GS, blocks = read_data(tmp_path / "data", 2)
    assert GS == cl["GS"] and len(energies) == len(cl["subspaces"])
    for k, e in cl["subspaces"]:
        assert np.array_equal(energies[k], e)
    assert set(blocks) == {("f", 0), ("f", 1), ("s", "n_d"), ("d", "A_d")}
    rx, rz = ini.read_coefficients(tmp_path / "data", 2)
    assert all(np.array_equal(a, b) for a, b in zip(rx, xi)) and all(np.array_equal(a, b) for a, b in zip(rz, ze))
