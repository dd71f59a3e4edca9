"""The C ABI library loads and exports every symbol include/smdp.h declares."""
import ctypes as C
import os
import re
import subprocess

import pytest

import smdp
from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "smdp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smdp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ["smdp_sssp", "smdp_bfs", "smdp_spmv_f32", "smdp_pagerank_f32", "smdp_color",
              "smdp_sssp_multi_source", "smdp_sssp_multi_graph", "smdp_autotune"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = smdp.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", smdp.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (smdp_\w+)", out))
    assert set(_declared()) <= exported


def test_exports_are_c_abi_only():
    out = subprocess.run(["nm", "-D", "--defined-only", smdp.LIB_PATH], capture_output=True, text=True).stdout
    # no torch types anywhere in the exported interface
    assert "torch" not in out and "at::" not in out


def test_strerror_codes():
    assert smdp.strerror(0) == "ok"
    assert "solution" in smdp.strerror(smdp.ERR_SOLUTION)
    assert smdp.strerror(-99) == "unknown error"


def test_create_without_device_fails_cleanly():
    # In a container without a GPU, context creation reports an error code.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    assert smdp.lib().smdp_create(0, C.byref(h)) == smdp.ERR_DEVICE
