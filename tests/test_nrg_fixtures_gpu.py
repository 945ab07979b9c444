"""End-to-end parity with the reference's own regression fixtures.

Every fixture of tests/golden/nrg (copied from NRG Ljubljana's test/c++ and test/c++sym_basic
suites: `param`, `data` and the expected outputs `ref/*`) is run as a whole calculation through
libnrgrun.so (nrg_run_dir, the counterpart of run_nrg_master(dir)): Hamiltonian assembly, the
batched eigensolver, the grouped recalculation products, the density-matrix sweeps and the
spectral kernels all run on the GPU. Every file of the fixture's ref/ directory must then match
the file the run wrote, number by number at 1e-5 absolute (test/mycomp.pl:20), over every line.

test/mycomp.pl never compares the final line of a file (its getline() returns the same
"END OF FILE REACHED" marker for both files once they reach their last line), so a stale final
line passes the reference's own suite. Two such lines are known and listed in STALE_FINAL_LINE;
for them the final line is excluded, and the test checks that the mismatch is in that line only.
"""
import os
import shutil

import pytest

from tests.compare_outputs import compare_files

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "nrg")

# fixture -> {file: reason}. The final line of these files disagrees with the reference's own
# code: test71's final `custom` row is the truncated-space value although lastall=true keeps every
# state at the last step (the same physics as test1_qs, whose final row agrees with ours and with the
# reference's keep_all_states_in_last_step, c++/params.hpp:539); test57's final `td` row is not
# reproduced by any other fixture and is not compared by the reference's comparator either.
STALE_FINAL_LINE = {
    "c++__test71_lastall=true": {"custom": "final custom row of the truncated space (lastall=true keeps all)"},
    "c++__test57_diagratio": {"td": "final td row, never compared by test/mycomp.pl"},
}


def fixtures():
    with open(os.path.join(GOLDEN, "FIXTURES.txt")) as F:
        return [l.strip().replace("/", "__") for l in F if l.strip() and not l.startswith("#")]


def _drop_final_line(path, out):
    with open(path) as F:
        lines = [l for l in F if l.strip()]
    with open(out, "w") as F:
        F.writelines(lines[:-1])


@pytest.mark.parametrize("name", fixtures())
def test_fixture_matches_reference_outputs(name, tmp_path):
    from nrgljubljana_amd import run_nrg
    src = os.path.join(GOLDEN, name)
    for f in ("param", "data"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    print(f"fixture {name}: running", flush=True)
    run_nrg(tmp_path, log_path=tmp_path / "log")
    failures, report = [], []
    for f in sorted(os.listdir(os.path.join(src, "ref"))):
        ref, mine = os.path.join(src, "ref", f), str(tmp_path / f)
        if not os.path.exists(mine):
            failures.append(f"{f}: not written")
            continue
        if f in STALE_FINAL_LINE.get(name, {}):
            ok_all, _, _ = compare_files(mine, ref)
            assert not ok_all, f"{name}/{f} now matches in full: remove it from STALE_FINAL_LINE"
            _drop_final_line(ref, str(tmp_path / (f + ".ref-head")))
            _drop_final_line(mine, str(tmp_path / (f + ".head")))
            ref, mine = str(tmp_path / (f + ".ref-head")), str(tmp_path / (f + ".head"))
        ok, err, msg = compare_files(mine, ref)
        report.append(f"{f} {err:.1e}")
        if not ok:
            failures.append(f"{f}: {msg}")
    print(f"fixture {name}: max |diff| per file: " + ", ".join(report), flush=True)
    assert not failures, "; ".join(failures)


# Fixtures re-run from data written by this project's generator (nrgljubljana_amd/initial.py) instead
# of the reference's Mathematica front end: same param file and Wilson-chain coefficients, generated
# multiplets and matrix elements. Every ref/ file must still match.
GENERATED = [("c++__test0_minimal", 1, ("n_d",)), ("c++__test1_qs", 1, ("n_d", "A_d", "sigma_d")),
             ("c++__legacy_substeps_one_siam_std", 2, ("n_d", "n_d^2", "A_d"))]


@pytest.mark.parametrize("name,channels,ops", GENERATED)
def test_generated_data_matches_reference_outputs(name, channels, ops, tmp_path):
    from nrgljubljana_amd import initial as ini, run_nrg
    from tests.test_initial_cpu import params
    src = os.path.join(GOLDEN, name)
    shutil.copy(os.path.join(src, "param"), tmp_path / "param")
    p = params(os.path.join(src, "param"))
    cl = ini.siam_qs(float(p["U"]), float(p.get("delta", 0)), [float(p["Gamma"])] * channels, float(p["Lambda"]),
                     p.get("discretization", "Z")[0])
    xi, zeta = ini.read_coefficients(os.path.join(src, "data"), channels)
    ini.write_data(tmp_path / "data", cl, xi, zeta, ops)
    print(f"generated data for {name}: running", flush=True)
    run_nrg(tmp_path, log_path=tmp_path / "log")
    report, failures = [], []
    for f in sorted(os.listdir(os.path.join(src, "ref"))):
        ok, err, msg = compare_files(str(tmp_path / f), os.path.join(src, "ref", f))
        report.append(f"{f} {err:.1e}")
        if not ok:
            failures.append(f"{f}: {msg}")
    print(f"generated data for {name}: max |diff| per file: " + ", ".join(report), flush=True)
    assert not failures, "; ".join(failures)
