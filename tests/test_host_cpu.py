"""CPU tests of the C++ host layer (nrgljubljana_amd/host/nrg): the container/helper unit tests of
tests/cpp/test_host_cpu.cpp, calc_ZnD / fdm_thermodynamics (c++/measurements.hpp:59-163) against
exact rational arithmetic, and the data-file reader (c++/read-input.hpp:23-135) on data files of the
reference's own tests (tests/golden/data).

The reference sums Z_n^D, Z_ZZG and the energy moments in 400-bit GMP floats and converts back with
mpf_get_d (truncation toward zero). Exact rationals over the same double summands stand in for the
400-bit sums (they differ from them by < 2^-390 relative), so the host layer's binary128 sums must
reproduce the reference's doubles bit for bit (E and C: within one ulp)."""
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_host_cpu")


def run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: build with tests/cpp/build.py")
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)


def test_unit_tests():
    r = run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def trunc(fr):
    """mpf_get_d: the double nearest to fr toward zero."""
    d = float(fr)
    if abs(Fraction(d)) > abs(fr):
        d = math.nextafter(d, 0.0)
    return d


def make_store(seed, T):
    rng = np.random.default_rng(seed)
    Nb, Ne, combs = 0, 8, 4
    subs = []
    tshift = 0.0
    for N in range(Nb, Ne):
        scale = 2.0 ** (-N / 2)
        nsub = int(rng.integers(2, 6))
        rels = [np.sort(rng.uniform(0.0, 6.0, int(rng.integers(1, 30)))) for _ in range(nsub)]
        shift = float(min(r[0] for r in rels))
        for r in rels:
            kept = int(rng.integers(0, len(r) + 1))
            subs.append(dict(N=N, mult=int(rng.integers(1, 4)), e=[float(x) for x in r], kept=kept,
                             last=int(N == Ne - 1), scale=scale, shift=shift, tshift=tshift))
        tshift += shift * scale * 0.5
    # absolute ground-state energy: lowest abs_T over all states
    absgs = min(((x - s["shift"]) * s["scale"] + s["tshift"]) for s in subs for x in s["e"])
    for s in subs:
        s["absgs"] = absgs
    return dict(T=T, combs=combs, Nb=Nb, Ne=Ne, GS=absgs * 0.5, subs=subs)


def write_store(st, path):
    with open(path, "w") as f:
        f.write(f"{st['T']!r} {st['combs']} {st['Nb']} {st['Ne']} {st['GS']!r} {len(st['subs'])}\n")
        for s in st["subs"]:
            f.write(f"{s['N']} {s['mult']} {len(s['e'])} {s['kept']} {s['last']} {s['scale']!r} {s['shift']!r} "
                    f"{s['tshift']!r} {s['absgs']!r} " + " ".join(repr(x) for x in s["e"]) + "\n")


def expected(st):
    T, combs, Nb, Ne = st["T"], st["combs"], st["Nb"], st["Ne"]
    ZG, ZN = {}, {}
    states = {N: [] for N in range(Nb, Ne)}
    for s in st["subs"]:
        lo = 0 if s["last"] else s["kept"]
        for x in s["e"][lo:]:
            abs_zero = (x - s["shift"]) * s["scale"]
            abs_T = abs_zero + s["tshift"]
            abs_G = abs_T - s["absgs"]
            states[s["N"]].append((float(s["mult"]), abs_zero, abs_T, abs_G))
    out = {}
    for N in range(Nb, Ne):
        ZG[N] = sum((Fraction(m * math.exp(-g / T)) for m, _, _, g in states[N]), Fraction(0))
        ZN[N] = sum((Fraction(m * math.exp(-z / T)) for m, z, _, _ in states[N]), Fraction(0))
    ZZG = trunc(sum((ZG[N] * combs ** (Ne - N - 1) for N in range(Nb, Ne)), Fraction(0)))
    wn, wnf = {}, {}
    for N in range(Nb, Ne):
        wnf[N] = float(combs ** (Ne - N - 1)) / ZZG
        wn[N] = wnf[N] * trunc(ZG[N])
        out[N] = (trunc(ZN[N]), wn[N], wnf[N])
    GS = st["GS"]
    E = E2 = Fraction(0)
    for N in range(Nb, Ne):
        if wn[N] > 1e-16:
            for m, z, t, _ in states[N]:
                w = Fraction(wn[N] * m * math.exp(-z / T)) / ZN[N]
                E += Fraction(t) * w
                E2 += Fraction(t) * Fraction(t) * w
    E_fdm = trunc(E)
    C_fdm = trunc(E2 - E * E) / T ** 2
    F_fdm = -math.log(ZZG) * T + GS
    out.update(ZZG=ZZG, Z_fdm=ZZG * math.exp(-GS / T), F_fdm=F_fdm, E_fdm=E_fdm, C_fdm=C_fdm,
               S_fdm=(E_fdm - F_fdm) / T)
    return out


def parse(stdout):
    res = {}
    for line in stdout.strip().splitlines():
        parts = line.split()
        if parts[0].isdigit():
            res[int(parts[0])] = tuple(float.fromhex(p) for p in parts[1:])
        else:
            res[parts[0]] = float.fromhex(parts[1])
    return res


def ulps(a, b):
    if a == b:
        return 0
    return abs(a - b) / math.ulp(max(abs(a), abs(b)))


@pytest.mark.parametrize("seed,T", [(1, 0.05), (2, 1e-3), (3, 0.7)])
def test_thermo_weights_match_exact_arithmetic(tmp_path, seed, T):
    st = make_store(seed, T)
    path = tmp_path / "store.txt"
    write_store(st, path)
    r = run("--thermo", str(path))
    assert r.returncode == 0, r.stdout + r.stderr
    got = parse(r.stdout)
    exp = expected(st)
    for N in range(st["Nb"], st["Ne"]):
        assert got[N] == exp[N], (N, got[N], exp[N])          # ZnDNd, w_n, w_n factor: bit for bit
    for k in ("ZZG", "Z_fdm", "F_fdm"):
        assert got[k] == exp[k], (k, got[k], exp[k])
    for k in ("E_fdm", "S_fdm"):
        assert ulps(got[k], exp[k]) <= 1, (k, got[k], exp[k])
    assert ulps(got["C_fdm"], exp["C_fdm"]) <= 1, (got["C_fdm"], exp["C_fdm"])
    assert abs(sum(got[N][1] for N in range(st["Nb"], st["Ne"])) - 1.0) < 1e-12   # sum rule of w_n


DATA = os.path.join(ROOT, "tests", "golden", "data")


def numeric_tokens(path):
    """Every number of a data file outside its '#' lines, in file order."""
    out = []
    with open(path) as f:
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            for tok in line.split():
                try:
                    out.append(float(tok))
                except ValueError:
                    pass
    return out


@pytest.mark.parametrize("name", sorted(f[:-4] for f in os.listdir(DATA) if f.endswith(".dat")))
def test_data_file_reader(tmp_path, name):
    """read_input -> write_input -> read_input gives the same data, and the rewritten file holds the
    same numbers as the original (count, exact sum and sum of squares): nothing skipped or misread."""
    src = os.path.join(DATA, name + ".dat")
    dst = tmp_path / "data"
    r = run("--data", src, str(dst))
    assert r.returncode == 0, r.stdout + r.stderr
    sym, channels, nmax, nsubs, nstates, perch, nops, ncoef, outcome = r.stdout.split()
    assert outcome == "same"
    with open(src) as f:
        text = f.read()
    assert sym == next(l.split()[-1] for l in text.splitlines() if "symtype" in l)
    head = next(l for l in text.splitlines() if l.strip() and not l.startswith("#")).split()
    assert (int(channels), int(nmax), int(nsubs)) == tuple(int(x) for x in head)
    a, b = numeric_tokens(src), numeric_tokens(dst)
    assert len(a) == len(b)
    assert math.fsum(a) == math.fsum(b) and math.fsum(x * x for x in a) == math.fsum(x * x for x in b)
    assert int(ncoef) >= 2 and int(nstates) >= int(nsubs)
