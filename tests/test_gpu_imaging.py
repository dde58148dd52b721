"""GPU parity of the imaging path through the C-ABI: gridder, FFTs and the full
imager against the FP64 oracle and the reference's gridder known-answer tests.
Tolerance (FP64 path): 1e-10 relative to the peak for images and grids."""
import json
import os

import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from fastimaging_amd.stp import STPError
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
KAT = json.load(open(os.path.join(GOLDEN, "gridder_kat.json")))["cases"]
KERNELS = {"TopHat": capi.KERNEL_TOPHAT, "Triangle": capi.KERNEL_TRIANGLE}
TOL = 1e-10


@pytest.mark.parametrize("case", KAT, ids=[c["name"] for c in KAT])
def test_gpu_gridder_kat(gpu, orc, case):
    uv = np.array(case["uv"], dtype=float)
    vis = np.array([complex(a, b) for a, b in case["vis"]])
    w = np.array(case["weights"], dtype=float)
    args = (KERNELS[case["kernel"]], case["param"], case["support"], case["image_size"], uv, vis, w)
    kw = dict(kernel_exact=case["kernel_exact"], oversampling=case["oversampling"], shift_uv=case["shift_uv"],
              halfplane=case["halfplane"], generate_beam=True)
    vg, sg, total = gpu.convolve_to_grid(*args, **kw)
    ovg, osg, ototal, _ = orc.convolve_to_grid(*args, **kw)
    assert np.allclose(vg, ovg, atol=TOL, rtol=0) and np.allclose(sg, osg, atol=TOL, rtol=0)
    assert abs(total - ototal) <= 1e-15 * abs(ototal)
    if "expected_vis_grid_real" in case:
        assert np.allclose(vg.real, np.array(case["expected_vis_grid_real"]), atol=TOL, rtol=0)
    else:
        exp = np.array(case["expected_vis_grid_real_after_column_shift"])
        assert np.allclose(np.roll(vg, case["column_shift"], axis=1).real, exp, atol=TOL, rtol=0)


@pytest.mark.parametrize("N,n,s,os_,exact,halfplane,shift", [
    (256, 5000, 3, 9, False, True, True),
    (512, 20000, 3, 8, False, True, True),
    (64, 3000, 2, 16, False, False, False),
    (128, 2000, 3, 1, True, True, True),
    (64, 500, 5, 4, False, True, True),
    (128, 1000, 7, 8, False, True, False),
    (128, 1500, 10, 4, False, True, True),  # 21 taps: the general tile kernel
])
def test_gpu_gridder_random(gpu, orc, N, n, s, os_, exact, halfplane, shift):
    rng = np.random.default_rng(N + n + s)
    uv = rng.uniform(-0.5 * N, 0.5 * N, size=(n, 2))
    uv[: n // 10, 1] = rng.uniform(-s - 1.5, s + 1.5, size=n // 10)  # near v = 0: conjugate copies
    vis = rng.normal(size=n) + 1j * rng.normal(size=n)
    w = rng.uniform(0.1, 2.0, size=n)
    args = (capi.KERNEL_GAUSSIANSINC, float(s), s, N, uv, vis, w)
    kw = dict(kernel_exact=exact, oversampling=os_, shift_uv=shift, halfplane=halfplane, generate_beam=True)
    vg, sg, total = gpu.convolve_to_grid(*args, **kw)
    ovg, osg, ototal, _ = orc.convolve_to_grid(*args, **kw)
    assert peak_rel_err(vg, ovg) < 1e-12
    assert peak_rel_err(sg, osg) < 1e-12
    assert abs(total - ototal) <= 1e-13 * ototal


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128, 256, 1024, 2048, 4096, 8192])
def test_gpu_fft_c2r(gpu, orc, n):
    rng = np.random.default_rng(n)
    g = rng.normal(size=(n // 2 + 1, n)) + 1j * rng.normal(size=(n // 2 + 1, n))
    out = gpu.fft_c2r(g)
    ref = orc.fft_c2r(g) if n <= 4096 else (np.fft.irfftn(g.T) * n * n).T
    assert peak_rel_err(out, ref) < 1e-12


@pytest.mark.parametrize("n,cols", [(16384, 64), (16384, 512)])
def test_gpu_fft_c2r_16384(gpu, n, cols):
    """The 2^14-point c2r along v (the reference's production image size,
    configs/reduce/fastimg_simple_config.json), on narrow grids."""
    rng = np.random.default_rng(n + cols)
    g = rng.normal(size=(n // 2 + 1, cols)) + 1j * rng.normal(size=(n // 2 + 1, cols))
    out = gpu.fft_c2r(g)
    ref = (np.fft.irfftn(g.T, s=(cols, n)) * n * cols).T
    assert peak_rel_err(out, ref) < 1e-12


def test_gpu_imager_16384(gpu, orc):
    """A 16384^2 image with the reference's production imager settings (PSWF support 3,
    oversampling 8, analytic grid correction, no beam) against the oracle."""
    snap = simdata.snapshot(16384, 0.5, 300_000, 20, seed=81)
    img, _, oimg, _ = _imaging_case(orc, gpu, snap, kernel_function=capi.KERNEL_PSWF, kernel_support=3,
                                    oversampling=8, generate_beam=False, gridding_correction=True, analytic_gcf=True)
    assert img.shape == (16384, 16384)
    e = peak_rel_err(img, oimg)
    print(f"N=16384: image err {e:.2e}", flush=True)
    assert e < TOL
    # and the source finder on the 2^28-pixel image, bit-exact against the oracle's
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    sf = gpu.source_find(img, sfp)
    lm, isl, _ = orc.source_find(img, sfp)
    assert np.array_equal(sf.label_map, lm)
    assert len(sf.islands) == len(isl) > 0


@pytest.mark.parametrize("n", [4, 16, 64, 256, 1024])
def test_gpu_fft_c2c(gpu, n):
    rng = np.random.default_rng(n + 1)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    assert peak_rel_err(gpu.fft_c2c(a, True), np.fft.fft2(a)) < 1e-12
    assert peak_rel_err(gpu.fft_c2c(a, False), np.fft.ifft2(a) * n * n) < 1e-12


def _imaging_case(orc, gpu, snap, **kw):
    N = snap["image_size"]
    pars = capi.imager_pars(N, snap["cell_arcsec"], **kw)
    img, beam, diag = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, return_diag=True)
    oimg, obeam, ototal, ongood = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    assert diag.num_good_vis == ongood
    assert abs(diag.sample_grid_total - ototal) <= 1e-13 * ototal
    return img, beam, oimg, obeam


def test_gpu_imager_c1(gpu, orc):
    """Workload c1: N=1024, 50k visibilities, GaussianSinc s=3, os 9, beam."""
    snap = simdata.workload_snapshot("c1")
    img, beam, oimg, obeam = _imaging_case(orc, gpu, snap)
    assert peak_rel_err(img, oimg) < TOL
    assert peak_rel_err(beam, obeam) < TOL


def test_gpu_imager_exact_and_gcf(gpu, orc):
    """Exact kernels with analytic grid correction (Gaussian) and padding 2 (crop)."""
    snap = simdata.snapshot(128, 2.0, 4000, 3, seed=77)
    img, beam, oimg, obeam = _imaging_case(orc, gpu, snap, kernel_function=capi.KERNEL_GAUSSIAN,
                                           kernel_exact=True, oversampling=1, gridding_correction=True,
                                           analytic_gcf=True)
    assert peak_rel_err(img, oimg) < 1e-9 and peak_rel_err(beam, obeam) < 1e-9
    snap = simdata.snapshot(128, 2.0, 4000, 3, seed=78)
    img, beam, oimg, obeam = _imaging_case(orc, gpu, snap, padding_factor=2.0, gridding_correction=True,
                                           analytic_gcf=False)
    assert img.shape == (128, 128)
    assert peak_rel_err(img, oimg) < TOL and peak_rel_err(beam, obeam) < TOL


def test_gpu_imager_pswf_defaults(gpu, orc):
    """The ImagerPars defaults (types.h:132-158): PSWF support 3, exact kernels,
    analytic grid correction, and the stp_python wrapper's defaults
    (stp_python.cpp:244-246) through the Python mirror."""
    from fastimaging_amd import stp_python
    snap = simdata.snapshot(256, 4.0, 20_000, 4, seed=79)
    pars = capi.ImagerPars()
    capi.load().stp_imager_pars_init(pars, 256, 4.0)
    assert pars.kernel_function == capi.KERNEL_PSWF and pars.kernel_exact == 1 and pars.gridding_correction == 1
    pars.generate_beam = 1
    img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars)
    oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    e_i, e_b = peak_rel_err(img, oimg), peak_rel_err(beam, obeam)
    print(f"PSWF exact + analytic gcf: image err {e_i:.2e}, beam err {e_b:.2e}", flush=True)
    assert e_i < TOL and e_b < TOL
    # the Python wrapper's defaults: PSWF exact, gridding correction from the
    # DFT of the kernel (analytic_gcf false), no beam
    img2, _ = stp_python.image_visibilities_wrapper(snap["vis"], snap["weights"], snap["uvw"], 256, 4.0)
    wpars = capi.imager_pars(256, 4.0, kernel_function=capi.KERNEL_PSWF, kernel_support=3, kernel_exact=True,
                             oversampling=8, generate_beam=False, gridding_correction=True, analytic_gcf=False)
    oimg2, _, _, _ = orc.image_visibilities(wpars, snap["vis"], snap["weights"], snap["uvw"])
    assert peak_rel_err(img2, oimg2) < TOL


@pytest.mark.parametrize("N,padding", [(512, 1.0), (256, 2.0)])
def test_gpu_imager_pswf_oversampled_analytic_gcf(gpu, orc, N, padding):
    """The reference's production imager settings (configs/reduce/fastimg_simple_config.json):
    PSWF support 3, oversampling 8, gridding correction with the analytic gcf, no beam;
    with padding 2 the image is cropped from the padded grid."""
    snap = simdata.snapshot(N, 2.0, 60_000, 8, seed=80)
    img, _, oimg, _ = _imaging_case(orc, gpu, snap, kernel_function=capi.KERNEL_PSWF, kernel_support=3,
                                    oversampling=8, generate_beam=False, gridding_correction=True,
                                    analytic_gcf=True, padding_factor=padding)
    assert img.shape == (N, N)
    assert peak_rel_err(img, oimg) < TOL


def test_gpu_imager_empty_and_errors(gpu, orc):
    """All visibilities out of bounds -> sample_grid_total 0 -> empty images (imager.h:355);
    malformed parameters -> status codes, not crashes."""
    N = 64
    uvw = np.zeros((10, 3), order="F")
    uvw[:, 0] = 1e12  # far outside the grid
    pars = capi.imager_pars(N, 1.0)
    img, beam = gpu.image_visibilities(np.ones(10, complex), np.ones(10), uvw, pars)
    assert img.size == 0
    bad = capi.imager_pars(96, 1.0)  # not a power of two
    with pytest.raises(STPError) as e:
        gpu.image_visibilities(np.ones(10, complex), np.ones(10), uvw, bad)
    assert e.value.status == 1
    ap = capi.AProjPars()
    ap.num_timesteps = 3
    with pytest.raises(STPError) as e:
        gpu.image_visibilities(np.ones(10, complex), np.ones(10), uvw, pars, a_proj=ap)
    assert e.value.status == 2


def test_gpu_imager_c2_full_size(gpu, orc):
    """Workload c2 at its full size: N=4096, 1M visibilities, 50 sources."""
    snap = simdata.workload_snapshot("c2")
    img, beam, oimg, obeam = _imaging_case(orc, gpu, snap)
    assert peak_rel_err(img, oimg) < TOL
    assert peak_rel_err(beam, obeam) < TOL
    # beam peak at the phase centre is 1 after normalisation
    assert abs(beam[0, 0] - 1.0) < 1e-12


_ORACLE_BEAMS = {}


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32C])
@pytest.mark.parametrize("N,padding,gcf", [(8, 1.0, False), (16, 1.0, True), (32, 2.0, True), (64, 1.0, False),
                                           (2048, 1.0, False), (8192, 1.0, False), (16384, 1.0, False)])
def test_gpu_beam_real_transform(gpu, orc, N, padding, gcf, precision):
    """Without W-projection the sampling grid is real and is transformed packed (two
    columns per complex element, half the bytes): every transform size from the
    smallest to 16384 (one and two columns of a pair per thread group), padding crop
    and grid correction, against the oracle's beam and the symmetry of a real even
    transform, B(y, x) = B(-y, -x) (padded size without crop)."""
    image = int(N / padding)
    snap = simdata.snapshot(image, 20.0 * 64 / max(image, 64), 3000 if N <= 64 else 20000, 3, seed=90 + N)
    pars = capi.imager_pars(image, snap["cell_arcsec"], padding_factor=padding, gridding_correction=gcf)
    img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars, precision=precision)
    key = (N, padding, gcf)
    if key not in _ORACLE_BEAMS:  # the oracle once per case, for both precisions
        _ORACLE_BEAMS.clear()
        _ORACLE_BEAMS[key] = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])[:2]
    oimg, obeam = _ORACLE_BEAMS[key]
    tol = TOL if precision == capi.PRECISION_FP64 else 1e-4
    e_b, e_i = peak_rel_err(beam, obeam), peak_rel_err(img, oimg)
    print(f"N={N} padding {padding} gcf {gcf} precision {precision}: beam err {e_b:.2e}, image err {e_i:.2e}", flush=True)
    assert e_b < tol and e_i < tol
    if padding == 1.0:
        refl = np.roll(np.roll(beam[::-1, ::-1], 1, axis=0), 1, axis=1)  # B(-y, -x)
        assert np.max(np.abs(refl - beam)) <= (1e-12 if precision == capi.PRECISION_FP64 else 1e-5) * np.max(np.abs(beam))
