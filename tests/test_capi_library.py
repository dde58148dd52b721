"""CPU checks of the C-ABI library: it loads without a GPU and exports every
entry point include/stp_capi.h declares, with matching struct sizes."""
import ctypes
import os
import re

import pytest

from fastimaging_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stp_capi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(stp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libstp_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # the ctypes signature table covers exactly the declared functions
    assert sorted(capi.EXPORTED_SYMBOLS) == names


def test_struct_layouts_match_header():
    # sizes implied by the C declarations (x86-64 alignment rules)
    assert ctypes.sizeof(capi.ImagerPars) == 56
    assert ctypes.sizeof(capi.WProjPars) == 40
    assert ctypes.sizeof(capi.SourceFindPars) == 56
    assert ctypes.sizeof(capi.Island) == 24 + 8 + 16 + 48
    assert ctypes.sizeof(capi.SourceFindStats) == 48
    assert ctypes.sizeof(capi.ImagerDiag) == 32


def test_ctx_create_without_device_fails_cleanly():
    """Without a HIP device the context constructor reports STP_ERR_DEVICE instead of crashing."""
    import subprocess
    import sys
    code = ("import ctypes,sys; sys.path.insert(0, %r); from fastimaging_amd import capi; "
            "lib=capi.load(); h=ctypes.c_void_p(); st=lib.stp_ctx_create(ctypes.byref(h), 0); print(st)") % ROOT
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() in ("3",)
