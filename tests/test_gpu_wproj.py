"""GPU parity of W-projection imaging (gridder.h:311-367,767-842;
aw_projection.cpp:19-210,366-474) against the FP64 oracle.  Bar: FP64 images and
beams within 1e-10 of the peak; the FP32 forms within 1e-4."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from fastimaging_amd.stp import STPError
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _wcase(gpu, orc, snap, N, cell, wp, precision=capi.PRECISION_FP64, **kw):
    pars = capi.imager_pars(N, cell, **kw)
    img, beam, diag = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, w_proj=wp,
                                             precision=precision, return_diag=True)
    oimg, obeam, ototal, ongood = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"], wproj=wp)
    assert diag.num_good_vis == ongood
    assert abs(diag.sample_grid_total - ototal) <= 1e-13 * ototal
    return img, beam, oimg, obeam


@pytest.mark.parametrize("planes,support,median,w_max", [(4, 8, False, 300.0), (7, 8, True, 300.0),
                                                         (1, 5, False, 300.0), (16, 12, False, 300.0),
                                                         (6, 14, False, 600.0), (5, 15, True, 600.0)])
def test_gpu_wproj_small(gpu, orc, planes, support, median, w_max):
    """N=256, 4000 visibilities with |w| <= w_max over the full field: plane assignment,
    W-kernel generation and gridding against the oracle.  At supports 14 and 15 the
    bounds check rejects visibilities at the grid edge, so the last plane's count
    (hence its mean |w|) stops before the end of the list it grids."""
    snap = simdata.snapshot(256, 20.0, 4000, 6, seed=300 + planes, field_frac=0.5, w_max=w_max)
    wp = capi.wproj_pars(planes, support, undersampling_opt=1, wplanes_median=median)
    img, beam, oimg, obeam = _wcase(gpu, orc, snap, 256, 20.0, wp)
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"wproj planes={planes} support={support} median={median}: image err {e_i:.3e}, beam err {e_b:.3e}")
    assert e_i < TOL and e_b < TOL


@pytest.mark.parametrize("trunc_perc,precision,tol", [(1.0, capi.PRECISION_FP64, TOL), (5.0, capi.PRECISION_FP64, TOL),
                                                       (1.0, capi.PRECISION_FP32C, 1e-4)])
def test_gpu_wproj_truncated_kernels(gpu, orc, trunc_perc, precision, tol):
    """kernel_trunc_perc > 0 (the reference's W-projection configs use 1 %): every plane
    grids with its own truncated support (aw_projection.cpp:197-209, gridder.h:756-758)."""
    snap = simdata.snapshot(256, 20.0, 5000, 6, seed=331, field_frac=0.5, w_max=600.0)
    wp = capi.wproj_pars(6, 14, undersampling_opt=1, kernel_trunc_perc=trunc_perc)
    img, beam, oimg, obeam = _wcase(gpu, orc, snap, 256, 20.0, wp, precision=precision)
    # truncation changes the image: the untruncated oracle differs by far more than tol
    full, _, _, _ = orc.image_visibilities(capi.imager_pars(256, 20.0), snap["vis"], snap["weights"], snap["uvw"],
                                           wproj=capi.wproj_pars(6, 14, undersampling_opt=1))
    assert peak_rel_err(oimg, full) > 100 * TOL
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"wproj truncated {trunc_perc}%: image err {e_i:.3e}, beam err {e_b:.3e}", flush=True)
    assert e_i < tol and e_b < tol


def test_gpu_wproj_conjugate_and_ties(gpu, orc):
    """Visibilities near v = 0 (conjugate copies with negated w, so conjugate kernels),
    equal |w| values straddling plane boundaries (stable |w| order), and weights != 1."""
    rng = np.random.default_rng(11)
    n = 3000
    N, cell = 128, 30.0
    cell_rad = cell / 3600.0 * np.pi / 180.0
    uvw = np.zeros((n, 3), order="F")
    uvw[:, 0] = rng.uniform(-0.4, 0.4, n) * N / (cell_rad * N)
    uvw[:, 1] = rng.uniform(-6.0, 6.0, n) / (cell_rad * N)     # |v| < support + 1 pixels
    uvw[:, 2] = rng.choice([-150.0, -75.0, 0.0, 75.0, 150.0, 149.9], n)
    vis = rng.normal(size=n) + 1j * rng.normal(size=n)
    w = rng.uniform(0.5, 2.0, n)
    snap = dict(uvw=uvw, vis=vis, weights=w)
    wp = capi.wproj_pars(5, 6)
    img, beam, oimg, obeam = _wcase(gpu, orc, snap, N, cell, wp)
    assert peak_rel_err(img, oimg) < TOL and peak_rel_err(beam, obeam) < TOL


@pytest.mark.parametrize("precision", [capi.PRECISION_FP32C, capi.PRECISION_FP32_TWOFLOAT])
def test_gpu_wproj_fp32(gpu, orc, precision):
    """The FP32 forms with W-projection: FP32 W-kernel caches and grids, FP64 kernel generation."""
    snap = simdata.snapshot(256, 20.0, 4000, 6, seed=321, field_frac=0.5, w_max=300.0)
    wp = capi.wproj_pars(4, 8)
    img, beam, oimg, obeam = _wcase(gpu, orc, snap, 256, 20.0, wp, precision=precision)
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"wproj precision {precision}: image err {e_i:.3e}, beam err {e_b:.3e}")
    assert e_i < 1e-4 and e_b < 1e-4


def test_gpu_wproj_errors(gpu):
    """Exact kernels and an unknown interpolation type are rejected with status codes."""
    snap = simdata.snapshot(64, 20.0, 100, 1, seed=5, w_max=10.0)
    wp = capi.wproj_pars(2, 4)
    with pytest.raises(STPError) as e:
        gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], capi.imager_pars(64, 20.0, kernel_exact=True),
                               w_proj=wp)
    assert e.value.status == capi.STATUS_CODES["STP_ERR_INVALID_ARGUMENT"]
    wp = capi.wproj_pars(2, 4, hankel_opt=True, interp_type=3)
    with pytest.raises(STPError) as e:
        gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], capi.imager_pars(64, 20.0), w_proj=wp)
    assert e.value.status == capi.STATUS_CODES["STP_ERR_INVALID_ARGUMENT"]


@pytest.mark.parametrize("proj_slice,interp,trunc", [(False, 0, 0.0), (False, 2, 0.0), (True, 0, 0.0), (True, 2, 0.0),
                                                     (False, 2, 2.0), (True, 1, 2.0)])
def test_gpu_wproj_hankel(gpu, orc, proj_slice, interp, trunc):
    """Hankel W-kernels (aw_projection.cpp:211-287,380-411): DHT of the radial profile or
    DFT of the projection, linear / cosine (linear) / cubic spline, with and without
    truncation, against the oracle at N=256; both FP64 and the FP32 path."""
    snap = simdata.snapshot(256, 20.0, 4000, 6, seed=341, field_frac=0.5, w_max=400.0)
    wp = capi.wproj_pars(5, 10, undersampling_opt=1, kernel_trunc_perc=trunc, hankel_opt=True,
                         hankel_proj_slice=proj_slice, interp_type=interp)
    img, beam, oimg, obeam = _wcase(gpu, orc, snap, 256, 20.0, wp)
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    img32, beam32, _, _ = _wcase(gpu, orc, snap, 256, 20.0, wp, precision=capi.PRECISION_FP32C)
    e32 = max(peak_rel_err(img32, oimg), peak_rel_err(beam32, obeam))
    # the Hankel kernels differ from the 2-D FFT ones: the images must differ by far more
    _, _, fimg, _ = _wcase(gpu, orc, snap, 256, 20.0, capi.wproj_pars(5, 10, undersampling_opt=1))
    print(f"hankel slice={proj_slice} interp={interp} trunc={trunc}: image err {e_i:.3e}, beam err {e_b:.3e}, "
          f"fp32 {e32:.3e}, vs 2-D kernels {peak_rel_err(oimg, fimg):.2e}", flush=True)
    assert peak_rel_err(oimg, fimg) > 100 * TOL
    assert e_i < TOL and e_b < TOL and e32 < 1e-4


def test_gpu_wproj_hankel_reference_config(gpu, orc):
    """The reference's Hankel configuration (configs/reduce-wproj/
    fastimg_config_wproj_HANKEL_size11.json: N=2048, 60" cells, Gaussian s3, os 8,
    support 127, projection slice, cubic spline, 1 % truncation) with 16 of its 128
    planes, against the oracle."""
    import os
    from fastimaging_amd.config import ConfigurationFile
    c = ConfigurationFile(os.path.join(os.path.dirname(__file__), "golden", "configs",
                                       "fastimg_config_wproj_HANKEL_size11.json"))
    c.w_proj.num_wplanes = 16
    N, cell = c.img_pars.image_size, c.img_pars.cell_size
    snap = simdata.snapshot(N, cell, 3000, 6, seed=351, field_frac=0.5, w_max=800.0)
    pars = c.img_pars
    img, beam, diag = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, w_proj=c.w_proj,
                                             return_diag=True)
    oimg, obeam, ototal, ongood = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"], wproj=c.w_proj)
    assert diag.num_good_vis == ongood > 0
    e = peak_rel_err(img, oimg)
    print(f"hankel reference config (16 planes, support 127): image err {e:.3e}", flush=True)
    assert e < TOL


C3_PARS = dict(kernel_support=3, oversampling=8)


@pytest.fixture(scope="module")
def c3_case(orc):
    """Workload c3 at its full size and its FP64 oracle image, computed once:
    N=4096, cell 2", 2M visibilities with |w| <= 1500, 40 sources over the full
    field, 32 planes, W-kernel support 32, oversampling 8."""
    snap = simdata.workload_snapshot("c3")
    wp = capi.wproj_pars(32, 32, undersampling_opt=1)
    pars = capi.imager_pars(4096, 2.0, **C3_PARS)
    oimg, obeam, ototal, ongood = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"], wproj=wp)
    return snap, wp, pars, oimg, obeam, ototal, ongood


@pytest.mark.parametrize("precision,tol", [(capi.PRECISION_FP64, TOL), (capi.PRECISION_FP32C, 1e-4),
                                           (capi.PRECISION_FP32_TWOFLOAT, 1e-4)])
def test_gpu_wproj_c3_full_size(gpu, c3_case, precision, tol):
    """c3 at full size on every precision path against the FP64 oracle: FP64 within
    1e-10 of the peak, the FP32 forms within 1e-4."""
    snap, wp, pars, oimg, obeam, ototal, ongood = c3_case
    img, beam, diag = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, w_proj=wp,
                                             precision=precision, return_diag=True)
    assert diag.num_good_vis == ongood
    assert abs(diag.sample_grid_total - ototal) <= 1e-13 * ototal
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"c3 precision {precision}: image err {e_i:.3e}, beam err {e_b:.3e}", flush=True)
    assert e_i < tol and e_b < tol
