"""Pins the oracle's source finder: bit-exact label maps from the reference's
labelling tests (golden fixture), source detection tests, moments, and an
imaging round trip (point source lands where the sign conventions put it)."""
import json
import math
import os

import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import fftshift, gaussian2d, noise

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LAB = json.load(open(os.path.join(GOLDEN, "labeling_kat.json")))


def _lab_pars(conn4):
    p = LAB["params"]
    return capi.sourcefind_pars(p["detection_n_sigma"], p["analysis_n_sigma"], rms_est=p["rms_est"],
                                find_negative_sources=True, sigma_clip_iters=5,
                                median_method=capi.MEDIAN_ZEROMEDIAN, ccl_4connectivity=conn4,
                                generate_labelmap=False, source_min_area=1)


@pytest.mark.parametrize("case", LAB["cases"], ids=[c["name"] for c in LAB["cases"]])
def test_labeling_kat(orc, case):
    img = np.array(case["image"], dtype=float)
    exp = np.array(case["expected_map_nonfftshift"])
    p = _lab_pars(case["ccl_4connectivity"])
    for sgn in (1, -1):
        lm, isl, st = orc.source_find(fftshift(sgn * img), p)
        assert np.array_equal(fftshift(lm), sgn * exp)
        assert len(isl) == case["expected_total_islands"]


@pytest.mark.parametrize("case", LAB["cases"], ids=[c["name"] for c in LAB["cases"]])
def test_labeling_reference_slab_scan(orc, case):
    """The literal restatement of the reference's slab scan with the 2-column slabs an 8x8
    image gets (ccl.h:137-147) reproduces the same maps."""
    img = fftshift(np.array(case["image"], dtype=float))
    exp = np.array(case["expected_map_nonfftshift"])
    lab, _, _ = orc.ccl(img, 3.0, -3.0, True, case["ccl_4connectivity"], sauf_slab_cols=2)
    assert np.array_equal(fftshift(lab), exp)


def test_ccl_matches_scipy_components(orc):
    """Component partition equals scipy.ndimage.label on the visual image (8- and 4-connected)."""
    from scipy import ndimage
    x = noise(64, 48, seed=7)
    for conn4 in (False, True):
        lab, npos, nneg = orc.ccl(x, 1.0, -1.0, True, conn4)
        vis = fftshift(lab)
        st = ndimage.generate_binary_structure(2, 1 if conn4 else 2)
        for sgn, cnt in ((1, npos), (-1, nneg)):
            ref, nref = ndimage.label(fftshift(sgn * x) > 1.0, structure=st)
            assert nref == cnt
            ours = np.where(sgn * vis > 0, sgn * vis, 0)
            # same partition: a bijection between label sets
            pairs = set(zip(ref[ref > 0].tolist(), ours[ref > 0].tolist()))
            assert len(pairs) == nref and len({a for a, _ in pairs}) == nref
            assert np.array_equal(ref > 0, ours > 0)


def test_ccl_ids_follow_scan_order(orc):
    """Ids are ranks of each component's first pixel along the scan: visual column major,
    memory row order inside a column (ccl.h:162-183,453-480)."""
    x = noise(32, 32, seed=8)
    lab, npos, _ = orc.ccl(x, 1.2, -1.2, True, False)
    rows, cols = x.shape
    seen = []
    for c in range(cols):
        mc = (c + cols // 2) % cols
        for r in range(rows):
            v = lab[r, mc]
            if v > 0 and v not in seen:
                seen.append(v)
    assert seen == list(range(1, npos + 1))


def _basic_detection(orc, find_negative, sign):
    """sourcefind_test_BasicSourceDetection.cpp:28-140 (NegativeSourceDetection with sign -1)."""
    ydim, xdim = 128, 64
    bright = gaussian2d(ydim, xdim, 48.24, 52.66, sign * 10.0, 1.5, 1.2, 1.0)
    faint = gaussian2d(ydim, xdim, 32, 64, sign * 3.5, 1.5, 1.2, 1.0)
    p = capi.sourcefind_pars(4, 3, rms_est=1.0, find_negative_sources=find_negative)
    img = bright.copy()
    _, isl0, _ = orc.source_find(fftshift(img), p)
    img += faint
    _, isl1, _ = orc.source_find(fftshift(img), p)
    img += faint
    _, isl2, _ = orc.source_find(fftshift(img), p)
    return isl0, isl1, isl2


def test_basic_source_detection(orc):
    isl0, isl1, isl2 = _basic_detection(orc, False, 1.0)
    assert len(isl0) == 1 and len(isl1) == 1 and len(isl2) == 2
    assert abs(isl0[0]["extremum_x_idx"] - 48.24) < 0.5 and abs(isl0[0]["extremum_y_idx"] - 52.66) < 0.5


def test_negative_source_detection(orc):
    isl0, isl1, isl2 = _basic_detection(orc, True, -1.0)
    assert len(isl0) == 1 and len(isl1) == 1 and len(isl2) == 2
    assert isl0[0]["sign"] == -1 and isl0[0]["label_idx"] == -1


def test_moments_of_symmetric_gaussian(orc):
    """Moments of an isolated Gaussian: barycentre equals the centre; fit axes ordered."""
    img = gaussian2d(96, 96, 40.0, 50.0, 20.0, 2.0, 2.0, 0.0)
    p = capi.sourcefind_pars(5, 3, rms_est=1.0, find_negative_sources=False)
    _, isl, _ = orc.source_find(fftshift(img), p)
    assert len(isl) == 1
    g = isl[0]["moments_fit"]
    assert abs(g[1] - 40.0) < 1e-9 and abs(g[2] - 50.0) < 1e-9
    assert g[3] >= g[4] > 0 and -math.pi / 2 < g[5] <= math.pi / 2


def test_label_map_detection_filter(orc):
    """generate_labelmap zeroes islands whose extremum does not pass detection (sourcefind.cpp:478-493),
    and min-area filters only the island list (:246)."""
    img = gaussian2d(64, 64, 20, 20, 10.0, 1.5, 1.5, 0.0) + gaussian2d(64, 64, 44, 44, 3.6, 1.5, 1.5, 0.0)
    p = capi.sourcefind_pars(4, 3, rms_est=1.0, find_negative_sources=False, generate_labelmap=True)
    lm, isl, st = orc.source_find(fftshift(img), p)
    assert st.num_labels_pos == 2 and st.num_valid_labels == 1 and len(isl) == 1
    assert set(np.unique(lm)) == {0, 1}
    p.generate_labelmap = 0
    lm2, _, _ = orc.source_find(fftshift(img), p)
    assert set(np.unique(lm2)) == {0, 1, 2}
    p.source_min_area = 10 ** 6
    _, isl3, st3 = orc.source_find(fftshift(img), p)
    assert len(isl3) == 0 and st3.num_valid_labels == 1


def test_imaging_point_source_position(orc):
    """A single point source (l, m) on a pixel centre images at un-shifted column l/cell, row m/cell
    with peak equal to its flux (weights 1, uv well inside the disk)."""
    N, cell = 256, 1.0
    snap = simdata.snapshot(N, cell, 20000, 1, seed=123, amp_range=(4.0, 4.0), noise_sigma=0.0)
    c = cell * simdata.ARCSEC
    l, m = 17 * c, -23 * c
    uvw = snap["uvw"]
    vis = 4.0 * np.exp(-2j * np.pi * (uvw[:, 0] * l + uvw[:, 1] * m))
    pars = capi.imager_pars(N, cell, kernel_function=capi.KERNEL_GAUSSIANSINC, kernel_support=3,
                            kernel_exact=False, oversampling=9, generate_beam=True)
    img, beam, total, ngood = orc.image_visibilities(pars, vis, snap["weights"], uvw)
    assert ngood == 20000
    r, cidx = np.unravel_index(np.argmax(img), img.shape)
    assert (r, cidx) == ((-23) % N, 17 % N)
    # without gridding correction the peak carries the image-domain response of the
    # anti-aliasing kernel (GaussianSinc passband ripple exceeds 1 near the centre)
    aa = orc.img_dom_kernel(capi.KERNEL_GAUSSIANSINC, 3, N)
    assert abs(img[r, cidx] - 4.0 * aa[N // 2 + 17] * aa[N // 2 + 23]) < 0.06
    assert abs(beam[0, 0] - 1.0) < 1e-12  # beam peak at the phase centre, normalised
