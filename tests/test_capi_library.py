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


STRUCTS = {
    "stp_imager_pars": "ImagerPars", "stp_kernel_spec": "KernelSpec", "stp_wproj_pars": "WProjPars",
    "stp_aproj_pars": "AProjPars", "stp_sourcefind_pars": "SourceFindPars", "stp_bounding_box": "BoundingBox",
    "stp_gaussian2d": "Gaussian2d", "stp_island": "Island", "stp_sourcefind_stats": "SourceFindStats",
    "stp_imager_diag": "ImagerDiag", "stp_server_cfg": "ServerCfg", "stp_job": "Job", "stp_job_result": "JobResult",
    "stp_batch_item": "BatchItem",
}


def test_struct_layouts_match_header(tmp_path):
    """sizeof and every field offset of the C structs (compiled from the header with gcc)
    equal the ctypes mirrors."""
    import subprocess
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "stp_capi.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(capi, pyname)
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    for cname, pyname in STRUCTS.items():
        cls = getattr(capi, pyname)
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


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
