"""Numeric comparison of NRG output files, the rule of the reference's test/mycomp.pl:
lines starting with '#' are skipped, lines are compared token by token, numeric tokens must agree
to 1e-5 absolute (mycomp.pl:20, ABSACCURACY), other tokens must be equal. Unlike mycomp.pl, the
number of compared lines must agree too. Returns (ok, max_abs_err, message)."""
import math
import re

ABSACCURACY = 1e-5
_num = re.compile(r"^[+-]?(?=\d|\.\d)\d*(\.\d*)?([Ee][+-]?\d+)?$")


def _lines(path):
    out = []
    with open(path) as F:
        for line in F:
            s = line.strip()
            if s.startswith("#"):
                continue
            out.append(s)
    while out and out[-1] == "":
        out.pop()
    return out


def compare_files(mine, ref, tol=ABSACCURACY):
    a, b = _lines(mine), _lines(ref)
    if len(a) != len(b):
        return False, math.inf, f"{len(a)} lines against {len(b)} in the reference"
    worst, bad = 0.0, []
    for k, (la, lb) in enumerate(zip(a, b)):
        if la == lb:
            continue
        ta, tb = la.split(), lb.split()
        if len(ta) != len(tb):
            bad.append(f"line {k + 1}: {len(ta)} tokens against {len(tb)}")
            continue
        for x, y in zip(ta, tb):
            if _num.match(x) and _num.match(y):
                d = abs(float(x) - float(y))
                worst = max(worst, d)
                if d > tol:
                    bad.append(f"line {k + 1}: {x} != {y}")
            elif x != y:
                bad.append(f"line {k + 1}: '{x}' != '{y}'")
    return not bad, worst, "; ".join(bad[:8])
