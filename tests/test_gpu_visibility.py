"""GPU parity of the model and residual visibilities (visibility.cpp:15-41, the
reduce driver's -d path reduce.cpp:113-122) against the oracle.  Bar: 1e-11
absolute per Jy of total sky flux (the reference's own test uses 1e-11 absolute,
visibility_test_ModelGeneration.cpp:15)."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from fastimaging_amd.stp import STPError
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu


def _sky(snap):
    return np.array([[l, m, a] for a, l, m in snap["sources"]])


def test_gpu_model_visibilities_c2_size(gpu, orc):
    """1M baselines with |w| <= 1500 and 50 sources over the field."""
    snap = simdata.snapshot(4096, 1.0, 1_000_000, 50, seed=simdata.seed_for(1, 41), amp_range=(0.5, 20.0),
                            log_amp=True, field_frac=0.5, w_max=1500.0, reference_model=True)
    sky = _sky(snap)
    v = gpu.generate_visibilities_from_local_skymodel(sky, snap["uvw"])
    ov = orc.model_visibilities(sky, snap["uvw"])
    err = float(np.max(np.abs(v - ov)))
    print(f"model visibilities: max abs err {err:.2e} (total flux {np.sum(sky[:, 2]):.1f} Jy)", flush=True)
    assert err < 1e-11 * np.sum(np.abs(sky[:, 2]))


def test_gpu_residual_device_and_image(gpu, orc):
    """d_vis -= model in place on the device, then the residual is imaged: equal to the
    oracle's residual and the oracle's image of it; the sources are gone from the image."""
    N, cell = 1024, 2.0
    snap = simdata.snapshot(N, cell, 200_000, 12, seed=simdata.seed_for(0, 42), field_frac=0.25,
                            reference_model=True)
    sky = _sky(snap)
    n = snap["weights"].size
    d_vis, d_w, d_uvw = gpu.malloc(16 * n), gpu.malloc(8 * n), gpu.malloc(24 * n)
    d_img = gpu.malloc(8 * N * N)
    try:
        gpu.htod(d_vis, np.ascontiguousarray(snap["vis"]))
        gpu.htod(d_w, np.ascontiguousarray(snap["weights"]))
        gpu.htod(d_uvw, np.asfortranarray(snap["uvw"]))
        gpu.model_visibilities_device(sky, n, d_uvw, d_vis, subtract=True, synchronize=True)
        res = np.empty(n, dtype=np.complex128)
        gpu.dtoh(res, d_vis)
        ores = snap["vis"] - orc.model_visibilities(sky, snap["uvw"])
        assert np.max(np.abs(res - ores)) < 1e-11 * np.sum(np.abs(sky[:, 2]))
        pars = capi.imager_pars(N, cell, generate_beam=False)
        gpu.image_visibilities_device(pars, n, d_vis, d_w, d_uvw, d_img, None, precision=capi.PRECISION_FP64,
                                      synchronize=True)
        img = np.empty((N, N), order="F")
        gpu.dtoh(img, d_img)
        oimg, _, _, _ = orc.image_visibilities(pars, ores, snap["weights"], snap["uvw"])
        assert peak_rel_err(img, oimg) < 1e-10
        full, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
        assert np.max(np.abs(img)) < 0.05 * np.max(np.abs(full))   # only the 0.1 Jy noise is left
    finally:
        for d in (d_vis, d_w, d_uvw, d_img):
            gpu.free(d)


def test_gpu_model_edge_cases(gpu):
    """An empty sky model gives a zero model and leaves the data as it is; a source
    outside the unit sphere is rejected."""
    uvw = np.asfortranarray(np.random.default_rng(1).uniform(-1e3, 1e3, (500, 3)))
    v = gpu.generate_visibilities_from_local_skymodel(np.zeros((0, 3)), uvw)
    assert np.all(v == 0)
    n = 500
    d_vis, d_uvw = gpu.malloc(16 * n), gpu.malloc(24 * n)
    try:
        data = np.random.default_rng(2).normal(size=n) + 1j * np.random.default_rng(3).normal(size=n)
        gpu.htod(d_vis, data)
        gpu.htod(d_uvw, uvw)
        gpu.model_visibilities_device(np.zeros((0, 3)), n, d_uvw, d_vis, subtract=True, synchronize=True)
        back = np.empty(n, dtype=np.complex128)
        gpu.dtoh(back, d_vis)
        assert np.array_equal(back, data)
    finally:
        gpu.free(d_vis)
        gpu.free(d_uvw)
    with pytest.raises(STPError) as e:
        gpu.generate_visibilities_from_local_skymodel([[0.9, 0.9, 1.0]], uvw)
    assert e.value.status == capi.STATUS_CODES["STP_ERR_INVALID_ARGUMENT"]
