"""Writes the QS and QSZ coefficient tables (1 and 2 channels) used by the host NRG driver.

The coefficients are data of NRG Ljubljana (GPL), generated there by Mathematica and kept in
c++/symmetry/coefnew/qs/*.dat and c++/symmetry/coefnew/qsz/*.dat of the reference, where they are
#include'd into C++ code. This script reads those files (it needs the reference checked out at
REF) and writes one plain-text table file per symmetry type and channel count, <SYM>-<n>ch.tab,
in this project's own format: a section per table, `@name`, then one row per line with the fields
separated by '|'. Numbers and expressions are kept as written (expressions in S(), SZ(), Power(),
sqrt() of the quantum-number variables q1, ss1, ssz1, ssp, sszp, ss); the driver evaluates them.

Run:  python nrgljubljana_amd/tables/make_tables.py [REF]   (REF defaults to /root/reference)
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# global singlet operators: name -> table of (i1, ip, ch, value) rows (recalc_global)
GLOBALS = {
    "QS": {"Qdiff": "qdiff", "Q1": "q1", "Q2": "q2", "Qtot": "qtot"},
    "QSZ": {"SZtot": "spinz", "Q1u": "q1u", "Q1d": "q1d"},
}

TABLES = {
    "QS": ["In2", "QN", "offdiag", "diag", "doubletp", "doubletm", "spinupa", "spindowna", "spinupb", "spindownb",
           "tripletp", "triplets", "tripletm", "offdiag-mix", "runghop"],
    "QSZ": ["In2", "QN", "offdiag", "diag", "doubletp", "doubletm", "spinupa", "spindowna", "spinupb", "spindownb",
            "tripletp", "triplets", "tripletm", "offdiag-UP", "offdiag-DOWN", "diag-UP", "diag-DOWN",
            "offdiag-mix", "runghop"],
}


def split_top(s):
    """Split at commas that are not inside parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    out.append(cur.strip())
    return out


def invar_args(s):
    m = re.fullmatch(r"Invar\((.*)\)", s.strip())
    if not m:
        raise ValueError(f"not an Invar: {s}")
    return split_top(m.group(1))


def convert(path):
    rows = []
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        m = re.fullmatch(r"(In|QN)\[(\d+)\]\s*=\s*(Invar\(.*\));", line)
        if m:
            rows.append([m.group(2)] + invar_args(m.group(3)))
            continue
        m = re.fullmatch(r"[A-Z_0-9]+\((.*)\);", line)
        if m:
            rows.append(split_top(m.group(1)))
            continue
        m = re.fullmatch(r"\{(.*)\},?", line)
        if m:
            f = split_top(m.group(1))
            row = []
            for x in f:
                row += invar_args(x) if x.startswith("Invar(") else [x]
            rows.append(row)
            continue
        raise ValueError(f"{path}: cannot parse line: {line}")
    return rows


def main(ref):
    for sym, names in TABLES.items():
        for ch in (1, 2):
            src = os.path.join(ref, "c++", "symmetry", "coefnew", sym.lower())
            out = [f"# {sym} symmetry, {ch} channel(s): coefficient tables of NRG Ljubljana (GPL),",
                   f"# from c++/symmetry/coefnew/{sym.lower()}/{sym.lower()}-{ch}ch-*.dat, converted by make_tables.py.",
                   "# Sections: In (ancestor quantum-number differences), QN (quantum numbers of the added shell",
                   "# states), offdiag (i|j|ch|factor), diag (i|ch|number), recalc tables i1|ip|IN1...|INp...|factor,",
                   "# f tables i1|ip|factor, global:<operator> tables i1|ip|ch|value."]
            for name in names:
                fn = os.path.join(src, f"{sym.lower()}-{ch}ch-{name}.dat")
                if not os.path.exists(fn):
                    continue
                out.append(f"@{'In' if name == 'In2' else name}")
                out += [" | ".join(r) for r in convert(fn)]
            for gname, fname in GLOBALS[sym].items():
                fn = os.path.join(src, f"{sym.lower()}-{ch}ch-{fname}.dat")
                if os.path.exists(fn):
                    out.append(f"@global:{gname}")
                    out += [" | ".join(r) for r in convert(fn)]
            with open(os.path.join(HERE, f"{sym}-{ch}ch.tab"), "w") as F:
                F.write("\n".join(out) + "\n")
            print(f"{sym}-{ch}ch.tab: {len(out)} lines")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
