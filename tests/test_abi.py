"""The C-ABI library loads without a GPU and exports every entry point include/nrgcore.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nrgcore.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nrg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["nrg_ctx_create", "nrg_ctx_destroy", "nrg_syev_batched", "nrg_gemm3_grouped", "nrg_fdm_bin"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from nrgljubljana_amd import core
    if not os.path.exists(core.LIB_PATH):
        pytest.fail("libnrgcore.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(core.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), f"{name} not exported"
    out = subprocess.check_output(["nm", "-D", "--defined-only", core.LIB_PATH]).decode()
    for name in declared():
        assert re.search(rf"\sT {name}$", out, re.M), f"{name} is not a defined text symbol"


def test_header_is_plain_c():
    subprocess.check_call(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", HEADER])


def test_create_without_device_fails_cleanly():
    from nrgljubljana_amd import core
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    L = core.lib()
    h = ctypes.c_void_p()
    r = L.nrg_ctx_create(ctypes.byref(h), 0)
    assert r != 0 and not h
    assert L.nrg_last_error()


RUN_HEADER = os.path.join(ROOT, "include", "nrgrun.h")


def run_declared():
    txt = re.sub(r"/\*.*?\*/", "", open(RUN_HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(nrg_[a-z0-9_]+)\s*\(", txt)))


def test_run_library_exports_every_declared_symbol():
    """libnrgrun.so (the whole-calculation entry of include/nrgrun.h) exports what its header declares."""
    from nrgljubljana_amd import core
    if not os.path.exists(core.RUN_LIB_PATH):
        pytest.fail("libnrgrun.so is not built (run __graft_entry__.build())")
    names = run_declared()
    assert names == sorted(core.RUN_EXPORTED)
    out = subprocess.check_output(["nm", "-D", "--defined-only", core.RUN_LIB_PATH]).decode()
    for name in names:
        assert re.search(rf"\sT {name}$", out, re.M), f"{name} is not a defined text symbol"
    subprocess.check_call(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", RUN_HEADER])


def test_run_without_device_fails_cleanly(tmp_path):
    """With no GPU the calculation stops with an error message, before writing any output."""
    from nrgljubljana_amd import core
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    src = os.path.join(ROOT, "tests", "golden", "nrg", "c++__test0_minimal")
    for f in ("param", "data"):
        with open(os.path.join(src, f)) as F, open(tmp_path / f, "w") as G:
            G.write(F.read())
    with pytest.raises(RuntimeError, match="nrg_run_dir failed"):
        core.run_nrg(tmp_path, log_path=tmp_path / "log")
    assert not (tmp_path / "td").exists()
