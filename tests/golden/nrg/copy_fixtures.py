"""Copies the reference's regression fixtures listed in FIXTURES.txt into this directory.

Each fixture is a directory of NRG Ljubljana's test suite (test/c++/<name> or
test/c++sym_basic/<name>, GPL test data): the input files `param` and `data` and the expected
outputs `ref/*` that the reference compares with test/compare.pl and test/mycomp.pl. Only those
files are copied (the copies of param and data that some ref/ directories hold are skipped). The
fixture `<suite>/<name>` becomes the directory `<suite>__<name>` here.

Run:  python tests/golden/nrg/copy_fixtures.py [REF]   (REF defaults to /root/reference)
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def fixtures():
    with open(os.path.join(HERE, "FIXTURES.txt")) as F:
        return [l.strip() for l in F if l.strip() and not l.startswith("#")]


def local_name(fixture):
    return fixture.replace("/", "__")


def main(ref):
    for fx in fixtures():
        src = os.path.join(ref, "test", fx)
        dst = os.path.join(HERE, local_name(fx))
        shutil.rmtree(dst, ignore_errors=True)
        os.makedirs(os.path.join(dst, "ref"))
        for f in ("param", "data"):
            shutil.copy(os.path.join(src, f), dst)
        for f in sorted(os.listdir(os.path.join(src, "ref"))):
            if f in ("param", "data") or not os.path.isfile(os.path.join(src, "ref", f)):
                continue
            shutil.copy(os.path.join(src, "ref", f), os.path.join(dst, "ref", f))
    print(f"{len(fixtures())} fixtures copied")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
