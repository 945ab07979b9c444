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
