"""GPU tests of the C++ host layer (tests/cpp/test_host_gpu.cpp): diagonalise / diagonalisations,
product / transform / rotate, recalc_general, the FDM density-matrix sweep and the FDM spectral step
through the nrgcore C ABI, each checked against the oracle and the reference's trace invariants."""
import os
import subprocess

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_host_gpu")

pytestmark = pytest.mark.gpu


def test_host_layer_on_gpu():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: build with tests/cpp/build.py")
    print("host layer: running tests/cpp/bin/test_host_gpu", flush=True)
    r = subprocess.run([BIN, "--blas=" + oracle.blas_path()], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
