"""The drivers' data formats (CPU): JSON configuration (src/auxiliary/load_json_config.h:29-197,
load_json_config.cpp:17-140), image-size rounding (src/reduce/common_functions.cpp:49-78),
NPZ arrays (src/auxiliary/load_data.h:19-149) and the island JSON document
(src/auxiliary/save_json_sf_output.cpp:9-52).  The configuration fixtures are the
reference's own files (configs/), copied under tests/golden/configs."""
import json
import os

import numpy as np
import pytest

from fastimaging_amd import capi
from fastimaging_amd.config import (ConfigurationFile, island_json, load_npy_complex_array, load_npy_double_array,
                                    set_image_sizes)

CFG = os.path.join(os.path.dirname(__file__), "golden", "configs")


def test_simple_config():
    c = ConfigurationFile(os.path.join(CFG, "fastimg_simple_config.json"))
    p = c.img_pars
    assert (p.image_size, p.padded_image_size, p.cell_size, p.padding_factor) == (16384, 16384, 0.5, 1.0)
    assert (p.kernel_function, p.kernel_support, p.kernel_exact, p.oversampling) == (capi.KERNEL_PSWF, 3, 0, 8)
    assert (p.generate_beam, p.gridding_correction, p.analytic_gcf, p.r_fft) == (0, 1, 1, 3)
    assert c.fft_wisdom_filename == "../wisdomfiles/WisdomFile_STP.fftw"
    assert c.w_proj.num_wplanes == 0 and c.a_proj.num_timesteps == 0
    assert (c.detection_n_sigma, c.analysis_n_sigma, c.find_negative_sources, c.estimate_rms) == (50.0, 50.0, True, 0.0)
    assert (c.sigma_clip_iters, c.median_method, c.gaussian_fitting) == (5, capi.MEDIAN_BINAPPROX, True)
    assert (c.ccl_4connectivity, c.generate_labelmap, c.source_min_area) == (False, False, 5)
    assert (c.s_ceres_diffmethod, c.s_ceres_solvertype) == ("AnalyticDiff_SingleResBlk", "LinearSearch_LBFGS")
    sp = c.sourcefind_pars()
    assert sp.detection_n_sigma == 50.0 and sp.median_method == capi.MEDIAN_BINAPPROX and sp.gaussian_fitting == 0


def test_wproj_and_aproj_config():
    c = ConfigurationFile(os.path.join(CFG, "fastimg_config_awproj_size11.json"))
    p, w, a = c.img_pars, c.w_proj, c.a_proj
    assert (p.image_size, p.cell_size, p.kernel_function) == (2048, 60.0, capi.KERNEL_GAUSSIAN)
    assert (w.num_wplanes, w.max_wpconv_support, w.undersampling_opt, w.kernel_trunc_perc) == (128, 127, 1, 1.0)
    assert (w.hankel_opt, w.hankel_proj_slice, w.wplanes_median, w.interp_type) == (0, 0, 0, 2)
    assert (a.num_timesteps, a.obs_dec, a.obs_ra) == (10, 47.339, 194.24)
    assert c.pbeam_coefs == [0, 0, 0, 0, 1, 0, 1, 0, 0]


def test_hankel_config():
    c = ConfigurationFile(os.path.join(CFG, "fastimg_config_wproj_HANKEL_size11.json"))
    p, w = c.img_pars, c.w_proj
    assert (p.image_size, p.cell_size, p.kernel_function, p.oversampling) == (2048, 60.0, capi.KERNEL_GAUSSIAN, 8)
    assert (w.num_wplanes, w.max_wpconv_support, w.kernel_trunc_perc) == (128, 127, 1.0)
    assert (w.hankel_opt, w.hankel_proj_slice, w.interp_type) == (1, 1, 2)


def test_pipeline_test_configs():
    g = ConfigurationFile(os.path.join(CFG, "fastimg_oversampling_gaussiansinc_config.json"))
    assert (g.img_pars.kernel_function, g.img_pars.oversampling, g.img_pars.generate_beam) == (capi.KERNEL_GAUSSIANSINC, 8, 1)
    assert g.img_pars.padded_image_size == 1024  # no padding_factor key: padded = image size
    assert g.median_method == capi.MEDIAN_BINMEDIAN and g.gaussian_fitting is False
    e = ConfigurationFile(os.path.join(CFG, "fastimg_exact_pswf_config.json"))
    assert e.img_pars.kernel_exact == 1 and e.img_pars.kernel_function == capi.KERNEL_PSWF


def test_config_defaults_and_errors():
    c = ConfigurationFile(document={})
    p = c.img_pars
    # the default-constructed ImagerPars (types.h:95-110) and ConfigurationFile members
    assert (p.image_size, p.kernel_function, p.kernel_exact, p.oversampling, p.generate_beam) == (0, capi.KERNEL_PSWF, 1, 0, 0)
    assert (c.gaussian_fitting, c.median_method, c.generate_labelmap) == (True, capi.MEDIAN_BINMEDIAN, False)
    assert c.w_proj.undersampling_opt == 1
    with pytest.raises(ValueError):
        ConfigurationFile(document={"imager_settings": {"kernel_function": "Bessel"}})
    with pytest.raises(ValueError):
        ConfigurationFile(document={"sourcefind_settings": {"median_method": "MEAN"}})


@pytest.mark.parametrize("size,pad,want", [(1024, 1.0, (1024, 1024)), (1022, 1.0, (1024, 1024)),
                                           (1000, 1.5, (1000, 1500)), (1001, 1.5, (1004, 1508))])
def test_set_image_sizes(size, pad, want):
    p = capi.imager_pars(size, 1.0, padding_factor=pad)
    notes = set_image_sizes(p)
    assert (p.image_size, p.padded_image_size) == want
    assert bool(notes) == (want != (size, int(size * pad)))


def test_npz_arrays(tmp_path):
    rng = np.random.default_rng(3)
    uvw = rng.normal(size=(7, 3))
    vis = (rng.normal(size=7) + 1j * rng.normal(size=7)).astype(np.complex64)
    w = rng.uniform(size=7).astype(np.float32)
    f = tmp_path / "data.npz"
    np.savez(f, uvw_lambda=uvw, uvw_f=np.asfortranarray(uvw), vis=vis, snr_weights=w, bad=np.arange(3))
    a, b = load_npy_double_array(str(f), "uvw_lambda"), load_npy_double_array(str(f), "uvw_f")
    assert a.shape == (7, 3) and np.array_equal(a, uvw) and np.array_equal(b, uvw) and a.flags.f_contiguous
    v = load_npy_complex_array(str(f), "vis")
    assert v.shape == (7, 1) and v.dtype == np.complex128 and np.array_equal(v[:, 0], vis.astype(np.complex128))
    assert load_npy_double_array(str(f), "snr_weights").dtype == np.float64
    with pytest.raises(ValueError):
        load_npy_double_array(str(f), "bad")
    with pytest.raises(ValueError):
        load_npy_complex_array(str(f), "uvw_lambda")
    np.save(tmp_path / "x.npy", uvw)
    assert np.array_equal(load_npy_double_array(str(tmp_path / "x.npy")), uvw)


def test_island_json_document():
    isl = [dict(label_idx=1, sign=1, extremum_y_idx=5, extremum_x_idx=6, num_samples=9, extremum_val=2.5,
                moments_fit=(2.5, 6.1, 5.2, 1.5, 1.1, 0.3)),
           dict(label_idx=-1, sign=-1, extremum_y_idx=9, extremum_x_idx=2, num_samples=7, extremum_val=-1.5,
                moments_fit=(-1.5, 2.0, 9.0, 1.2, 1.0, -0.2))]
    fits = [dict(leastsq_fit=(2.6, 6.0, 5.1, 1.4, 1.0, 0.31)), dict(leastsq_fit=(-1.6, 2.1, 9.1, 1.3, 1.1, -0.1))]
    doc = json.loads(json.dumps(island_json(isl, fits)))
    assert list(doc) == ["Islands"] and len(doc["Islands"]) == 2
    i0 = doc["Islands"][0]
    assert list(i0) == ["sign", "val", "x_idx", "y_idx", "MomentsFit", "GaussianFit"]
    assert (i0["sign"], i0["val"], i0["x_idx"], i0["y_idx"]) == (1, 2.5, 6, 5)
    assert list(i0["MomentsFit"]) == ["amplitude", "x_centre", "y_centre", "semimajor", "semiminor", "theta"]
    assert i0["GaussianFit"]["theta"] == 0.31
    assert "GaussianFit" not in island_json(isl)["Islands"][1]
