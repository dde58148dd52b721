"""The reduce driver (src/reduce/reduce.cpp:74-173) end to end on the GPU against the
oracle's pipeline on the same files: configuration JSON and NPZ in, residual
visibilities (-d), imaging, source finding and the island fit, island JSON and
label-map NPZ out."""
import json

import numpy as np
import pytest

from fastimaging_amd import capi, reduce
from tests import simdata

pytestmark = pytest.mark.gpu


def _write_inputs(tmp_path, n_sources_kept):
    """A 512^2 snapshot whose NPZ skymodel holds the first sources of the sky, so the
    residual keeps the others."""
    N, cell = 512, 2.0
    snap = simdata.snapshot(N, cell, 30_000, 6, seed=simdata.seed_for(1, 77))
    # skymodel rows (l, m, flux) whose model visibilities (visibility.cpp:15-28,
    # flux n e^{+2 pi i (u l + v m)} at w = 0) equal the snapshot's A e^{-2 pi i (u l + v m)}
    src = np.asarray(snap["sources"], dtype=np.float64)  # (A, l, m)
    n = np.sqrt(1.0 - src[:, 1] ** 2 - src[:, 2] ** 2)
    sky = np.column_stack([-src[:, 1], -src[:, 2], src[:, 0] / n])
    cfg = {"imager_settings": {"image_size_pix": N, "cell_size_arcsec": cell, "kernel_function": "GaussianSinc",
                               "kernel_support": 3, "kernel_exact": False, "oversampling": 9, "generate_beam": False,
                               "gridding_correction": False, "analytic_gcf": False},
           "sourcefind_settings": {"sourcefind_detection": 5.0, "sourcefind_analysis": 3.0,
                                   "find_negative_sources": True, "median_method": "BINMEDIAN",
                                   "gaussian_fitting": True, "generate_labelmap": True, "source_min_area": 5}}
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    np.savez(tmp_path / "data.npz", uvw_lambda=snap["uvw"], vis=snap["vis"], snr_weights=snap["weights"],
             skymodel=sky[:n_sources_kept])
    return snap, sky, N, cell


@pytest.mark.parametrize("diff", [False, True])
def test_reduce_matches_oracle(gpu, orc, tmp_path, diff):
    snap, sky, N, cell = _write_inputs(tmp_path, 3)
    r = reduce.run(str(tmp_path / "cfg.json"), str(tmp_path / "data.npz"), diff=diff,
                   out_json=str(tmp_path / "isl.json"), out_npz=str(tmp_path / "lab.npz"), ctx=gpu)
    vis = snap["vis"] - orc.model_visibilities(sky[:3], snap["uvw"]) if diff else snap["vis"]
    pars = capi.imager_pars(N, cell, generate_beam=False)
    oimg, _, _, _ = orc.image_visibilities(pars, vis, snap["weights"], snap["uvw"])
    assert np.max(np.abs(r["image"] - oimg)) <= 1e-10 * np.max(np.abs(oimg))
    sfp = r["config"].sourcefind_pars()
    lm, isl, _ = orc.source_find(r["image"], sfp)
    assert np.array_equal(r["label_map"], lm)
    key = lambda i: (i["label_idx"], i["sign"], i["extremum_x_idx"], i["extremum_y_idx"], i["num_samples"])
    assert [key(i) for i in r["islands"]] == [key(i) for i in isl]
    ofits = orc.fit_islands(r["image"], lm, isl)
    for i, a, b in zip(isl, r["fits"], ofits):
        assert a["bounding_box"] == b["bounding_box"]
        if i["num_samples"] >= 12 and b["status"] == 0:
            assert np.allclose(a["leastsq_fit"][:5], b["leastsq_fit"][:5], rtol=1e-6, atol=1e-6)
    doc = json.loads((tmp_path / "isl.json").read_text())
    assert len(doc["Islands"]) == len(isl)
    assert all("GaussianFit" in d for d in doc["Islands"])
    with np.load(tmp_path / "lab.npz", allow_pickle=False) as z:
        assert np.array_equal(z["label_map"], lm)
    # the residual removes the modelled sources: the brightest of them leaves no island there
    if diff:
        l, m = -sky[np.argmax(sky[:3, 2]), :2]
        x = int(round(N / 2 + l / (cell * np.pi / 180 / 3600))) % N
        y = int(round(N / 2 + m / (cell * np.pi / 180 / 3600))) % N
        assert not any(abs(i["extremum_x_idx"] - x) <= 1 and abs(i["extremum_y_idx"] - y) <= 1 for i in isl)
    print(f"reduce diff={diff}: {len(isl)} islands, stage ms "
          + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in r["timings"].items()), flush=True)


def test_reduce_cli(tmp_path, capsys):
    _write_inputs(tmp_path, 2)
    rc = reduce.main([str(tmp_path / "cfg.json"), str(tmp_path / "data.npz"), "-d", "-s", str(tmp_path / "o.json"),
                      "-o", str(tmp_path / "o.npz")])
    assert rc == 0
    out = capsys.readouterr().out
    assert "Number of detected sources:" in out and "imager:" in out
    assert (tmp_path / "o.json").exists() and (tmp_path / "o.npz").exists()
