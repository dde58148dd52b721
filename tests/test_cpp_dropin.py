"""The C++ drop-in (include/stp/stp.h): a caller written against the reference's
stp:: API (tests/cpp/dropin_caller.cpp) compiles and links against libstp_hip.so
(CPU test), and on the GPU its ImageVisibilities / image_visibilities /
SourceFindImage / generate_visibilities_from_local_skymodel results equal the
oracle's (images within 1e-10 of the peak, label map and islands bit-exact,
model visibilities within 1e-11 per Jy)."""
import os
import subprocess

import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import peak_rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_caller.cpp")
LIBDIR = os.path.join(ROOT, "fastimaging_amd", "lib")


def _build(tmp_path):
    exe = str(tmp_path / "dropin_caller")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lstp_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_header_compiles_and_links(tmp_path):
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libstp_hip.so is not built (run __graft_entry__.build())")
    exe = _build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_caller_matches_oracle(tmp_path, orc):
    exe = _build(tmp_path)
    snap = simdata.workload_snapshot("c1")
    N, cell = 1024, 1.0
    n = snap["weights"].size
    sky = np.array([[l, m, a] for a, l, m in snap["sources"]])
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([n, N], dtype=np.int64).tofile(f)
        np.array([cell], dtype=np.float64).tofile(f)
        np.array([sky.shape[0]], dtype=np.int64).tofile(f)
        np.asfortranarray(snap["uvw"]).ravel(order="F").tofile(f)
        np.ascontiguousarray(snap["vis"]).view(np.float64).tofile(f)
        np.ascontiguousarray(snap["weights"]).tofile(f)
        np.asfortranarray(sky).ravel(order="F").tofile(f)
    out = str(tmp_path / "out_")
    r = subprocess.run([exe, str(inp), out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rd = lambda name, dt: np.fromfile(out + name, dtype=dt)
    img = rd("image.f64", np.float64).reshape((N, N), order="F")
    beam = rd("beam.f64", np.float64).reshape((N, N), order="F")
    img2 = rd("image2.f64", np.float64).reshape((N, N), order="F")
    pars = capi.imager_pars(N, cell)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    assert peak_rel_err(img, oimg) < 1e-10 and peak_rel_err(beam, obeam) < 1e-10
    assert peak_rel_err(img2, oimg) < 1e-10
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    lm, isl, st = orc.source_find(img, sfp)
    labels = rd("labels.i32", np.int32).reshape((N, N), order="F")
    assert np.array_equal(labels, lm)
    got = rd("islands.f64", np.float64)
    rec = got[:-2].reshape(-1, 6)
    assert rec.shape[0] == len(isl) > 0
    for row, i in zip(rec, isl):
        assert tuple(row) == (i["label_idx"], i["sign"], i["extremum_y_idx"], i["extremum_x_idx"], i["num_samples"],
                              i["extremum_val"])
    assert got[-2] == st.rms_est and got[-1] == st.bg_level
    model = rd("model.f64", np.float64).view(np.complex128)
    omodel = orc.model_visibilities(sky, snap["uvw"])
    assert np.max(np.abs(model - omodel)) < 1e-11 * np.sum(np.abs(sky[:, 2]))
