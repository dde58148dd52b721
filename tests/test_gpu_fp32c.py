"""GPU tests of the FP32 paths (STP_PRECISION_FP32C: FP64 tile sums;
STP_PRECISION_FP32_TWOFLOAT: two-float tile sums) against the FP64 oracle.  Bar: images and beams within 1e-4 of the peak (the FP32 form's
tolerance); islands of the FP32 image identical in sign and extremum pixel to
the oracle's, fluxes within 1e-4 of the peak.  The FP32 source finder itself
(FP32 image in, FP64 accumulations) is bit-exact against the oracle on the
same FP32 values."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import islands_agree_banded, compare, peak_rel_err

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4


def _device_pipeline(gpu, snap, pars, sfp, precision):
    """Imaging + source finding with every buffer in device memory."""
    n = snap["vis"].size
    N = pars.image_size
    dt = np.float64 if precision == capi.PRECISION_FP64 else np.float32
    osz = np.dtype(dt).itemsize
    bufs = [gpu.malloc(16 * n), gpu.malloc(8 * n), gpu.malloc(24 * n), gpu.malloc(osz * N * N),
            gpu.malloc(osz * N * N), gpu.malloc(4 * N * N)]
    d_vis, d_w, d_uvw, d_img, d_beam, d_lab = bufs
    try:
        gpu.htod(d_vis, np.ascontiguousarray(snap["vis"]))
        gpu.htod(d_w, np.ascontiguousarray(snap["weights"]))
        gpu.htod(d_uvw, np.asfortranarray(snap["uvw"]))
        gpu.image_visibilities_device(pars, n, d_vis, d_w, d_uvw, d_img, d_beam, precision=precision, synchronize=True)
        img = np.empty((N, N), order="F", dtype=dt)
        beam = np.empty((N, N), order="F", dtype=dt)
        gpu.dtoh(img, d_img)
        gpu.dtoh(beam, d_beam)
        res = None
        if sfp is not None:
            isl, st = gpu.source_find_device(d_img, N, N, sfp, d_label_map=d_lab, precision=precision)
            res = (list(isl), st)
        return img, beam, res
    finally:
        for b in bufs:
            gpu.free(b)


@pytest.mark.parametrize("precision", [capi.PRECISION_FP32C, capi.PRECISION_FP32_TWOFLOAT])
def test_fp32c_imager_c1(gpu, orc, precision):
    """Workload c1 (N=1024, 50k visibilities) in both FP32 forms: image and beam within 1e-4 of the peak."""
    snap = simdata.workload_snapshot("c1")
    pars = capi.imager_pars(1024, snap["cell_arcsec"])
    img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, precision=precision)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    e_img, e_beam = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"c1 precision {precision}: image err {e_img:.3e}, beam err {e_beam:.3e}")
    assert e_img < FP32_TOL and e_beam < FP32_TOL
    # FP32 arithmetic, not FP64: the error is above the FP64 bar
    assert e_img > 1e-10


@pytest.mark.parametrize("support,precision", [(7, capi.PRECISION_FP32C), (7, capi.PRECISION_FP32_TWOFLOAT),
                                               (10, capi.PRECISION_FP32C), (10, capi.PRECISION_FP32_TWOFLOAT)])
def test_fp32_wide_kernels(gpu, orc, support, precision):
    """Oversampled kernels of 15 taps (the specialised tile kernel's 16-row form) and of
    21 taps (the general tile kernel) in both FP32 forms, within 1e-4 of the peak."""
    snap = simdata.snapshot(256, 4.0, 20000, 5, seed=90 + support)
    pars = capi.imager_pars(256, 4.0, kernel_support=support, oversampling=8)
    img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, precision=precision)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    assert peak_rel_err(img, oimg) < FP32_TOL and peak_rel_err(beam, obeam) < FP32_TOL


@pytest.mark.parametrize("exact,precision", [(False, capi.PRECISION_FP32C), (True, capi.PRECISION_FP32C),
                                             (True, capi.PRECISION_FP32_TWOFLOAT)])
def test_fp32c_exact_padding_gcf(gpu, orc, exact, precision):
    """Exact kernels, padding 2 (crop) and grid correction in the FP32 forms."""
    snap = simdata.snapshot(128, 2.0, 4000, 3, seed=91)
    kw = dict(kernel_function=capi.KERNEL_GAUSSIAN, kernel_exact=exact, oversampling=1 if exact else 8,
              gridding_correction=True, analytic_gcf=True, padding_factor=2.0)
    pars = capi.imager_pars(128, 2.0, **kw)
    img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, precision=precision)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    assert peak_rel_err(img, oimg) < FP32_TOL and peak_rel_err(beam, obeam) < FP32_TOL


@pytest.mark.parametrize("precision", [capi.PRECISION_FP32C, capi.PRECISION_FP32_TWOFLOAT])
def test_fp32_compensated_accumulation(gpu, orc, precision):
    """200k identical visibilities on one uv point: every footprint cell sums 200k
    FP32 terms of 0.1 w K, where plain FP32 accumulation drifts by ~1e-3 relative;
    FP64 and two-float tile sums both keep the image at the FP32 rounding level of
    the kernel values and the FFT (~1e-6 of the peak)."""
    N, cell = 64, 10.0
    n = 200_000
    cell_rad = cell / 3600.0 * np.pi / 180.0
    u_px, v_px = 10.3, 20.7
    uvw = np.zeros((n, 3), order="F")
    uvw[:, 0] = u_px / (cell_rad * N)
    uvw[:, 1] = v_px / (cell_rad * N)
    vis = np.full(n, 0.1 + 0.05j)
    w = np.ones(n)
    pars = capi.imager_pars(N, cell)
    img, beam = gpu.image_visibilities(vis, w, uvw, pars, precision=precision)
    oimg, obeam, _, _ = orc.image_visibilities(pars, vis, w, uvw)
    e_img, e_beam = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"compensated accumulation (precision {precision}): image err {e_img:.3e}, beam err {e_beam:.3e}")
    assert e_img < 1e-5 and e_beam < 1e-5


def test_fp32_sourcefinder_bit_exact(gpu, orc):
    """The source finder on an FP32 image (FP64 median, RMS and moments accumulations)
    is bit-exact against the oracle run on the same FP32 values."""
    snap = simdata.workload_snapshot("c1")
    pars = capi.imager_pars(1024, snap["cell_arcsec"])
    oimg, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    img32 = np.asfortranarray(oimg.astype(np.float32))
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    N = 1024
    d_img, d_lab = gpu.malloc(4 * N * N), gpu.malloc(4 * N * N)
    try:
        gpu.htod(d_img, img32)
        isl, st = gpu.source_find_device(d_img, N, N, sfp, d_label_map=d_lab, precision=capi.PRECISION_FP32C)
        lab = np.empty((N, N), order="F", dtype=np.int32)
        gpu.dtoh(lab, d_lab)
    finally:
        gpu.free(d_img)
        gpu.free(d_lab)

    class R:
        pass
    r = R()
    r.label_map, r.islands, r.stats = lab, list(isl), st
    compare(r, orc.source_find(np.asfortranarray(img32.astype(np.float64)), sfp))


@pytest.mark.parametrize("precision", [capi.PRECISION_FP32C, capi.PRECISION_FP32_TWOFLOAT])
def test_fp32c_pipeline_c2(gpu, orc, precision):
    """Workload c2 at full size (N=4096, 1M visibilities) through the FP32 device pipeline:
    image within 1e-4 of the peak, and against the islands the oracle finds on the FP64
    oracle image: every island whose extremum clears its detection threshold by more
    than the guard band (image error plus the threshold change) is found on both sides
    with the same sign, extremum pixel and extremum value within the band."""
    snap = simdata.workload_snapshot("c2")
    pars = capi.imager_pars(4096, snap["cell_arcsec"])
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    img, beam, (isl, st) = _device_pipeline(gpu, snap, pars, sfp, precision)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    peak = np.max(np.abs(oimg))
    e_img, e_beam = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    _, oisl, ost = orc.source_find(oimg, sfp)
    assert e_img < FP32_TOL and e_beam < FP32_TOL
    assert abs(st.rms_est - ost.rms_est) < FP32_TOL * peak
    # guard band: a pixel's distance to a detection threshold moves by at most the image
    # error plus the change of the threshold (background + 5 rms) between the two images
    band = e_img * peak + abs(st.bg_level - ost.bg_level) + sfp.detection_n_sigma * abs(st.rms_est - ost.rms_est)
    det_pos = ost.bg_level + sfp.detection_n_sigma * ost.rms_est
    det_neg = ost.bg_level - sfp.detection_n_sigma * ost.rms_est
    strict, banded = islands_agree_banded(list(isl), list(oisl), det_pos, det_neg, band)
    key = lambda i: (i["sign"], i["extremum_x_idx"], i["extremum_y_idx"])
    same = [key(i) for i in isl] == [key(i) for i in oisl]
    print(f"c2 precision {precision}: image err {e_img:.3e}, beam err {e_beam:.3e}, islands {len(isl)} vs {len(oisl)}, "
          f"{strict} beyond the guard band {band / peak:.2e} of the peak, {banded} inside it, lists identical: {same}",
          flush=True)
    assert strict > 0.9 * len(oisl)
