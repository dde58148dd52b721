"""GPU parity of the source finder through the C-ABI.  Bar: label maps, label
counts, island ids/positions/areas and the statistics (mean, sigma, median,
RMS) bit-exact against the oracle on the same image; the moments fit within the
tolerances of fits_agree()."""
import itertools
import json
import os

import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import ISLAND_EXACT, compare, fftshift, fits_agree, gaussian2d, noise

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LAB = json.load(open(os.path.join(GOLDEN, "labeling_kat.json")))
@pytest.mark.parametrize("case", LAB["cases"], ids=[c["name"] for c in LAB["cases"]])
def test_gpu_labeling_kat(gpu, case):
    img = np.array(case["image"], dtype=float)
    exp = np.array(case["expected_map_nonfftshift"])
    p = LAB["params"]
    pars = capi.sourcefind_pars(p["detection_n_sigma"], p["analysis_n_sigma"], rms_est=p["rms_est"],
                                median_method=capi.MEDIAN_ZEROMEDIAN, ccl_4connectivity=case["ccl_4connectivity"],
                                generate_labelmap=False, source_min_area=1)
    for sgn in (1, -1):
        r = gpu.source_find(fftshift(sgn * img), pars)
        assert np.array_equal(fftshift(r.label_map), sgn * exp)
        assert len(r.islands) == case["expected_total_islands"]


def _field(rows, cols, seed, n_src=12, amp=(3.0, 30.0)):
    rng = np.random.default_rng(seed)
    img = noise(rows, cols, sigma=1.0, seed=seed)
    for _ in range(n_src):
        a = rng.uniform(*amp) * rng.choice([-1.0, 1.0])
        img += gaussian2d(rows, cols, rng.uniform(0, cols), rng.uniform(0, rows), a, rng.uniform(0.8, 3.0),
                          rng.uniform(0.8, 3.0), rng.uniform(0, np.pi))
    return np.asfortranarray(img)


MEDIANS = [capi.MEDIAN_BINMEDIAN, capi.MEDIAN_BINAPPROX, capi.MEDIAN_NTHELEMENT, capi.MEDIAN_ZEROMEDIAN]


@pytest.mark.parametrize("median,conn4,find_neg,gen_map", list(itertools.product(
    MEDIANS, [False, True], [True, False], [True, False]))[::3])
def test_gpu_sourcefind_modes(gpu, orc, median, conn4, find_neg, gen_map):
    img = _field(256, 192, seed=median * 10 + conn4)
    pars = capi.sourcefind_pars(5.0, 3.0, median_method=median, ccl_4connectivity=conn4,
                                find_negative_sources=find_neg, generate_labelmap=gen_map)
    compare(gpu.source_find(img, pars), orc.source_find(img, pars))


@pytest.mark.parametrize("rows,cols,seed", [(64, 64, 1), (130, 258, 2), (512, 512, 3), (1024, 1024, 4)])
def test_gpu_sourcefind_sizes(gpu, orc, rows, cols, seed):
    img = _field(rows, cols, seed=seed, n_src=max(3, rows * cols // 20000))
    pars = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    compare(gpu.source_find(img, pars), orc.source_find(img, pars))


def test_gpu_sourcefind_given_rms_and_iters(gpu, orc):
    img = _field(200, 200, seed=11)
    for rms, iters in ((0.9, 5), (0.0, 1), (0.0, 0), (0.0, 12)):
        pars = capi.sourcefind_pars(5.0, 3.0, rms_est=rms, sigma_clip_iters=iters, median_method=capi.MEDIAN_BINAPPROX)
        compare(gpu.source_find(img, pars), orc.source_find(img, pars))


def _clip_bounds(img, iters, num_sigma=3.0):
    """Bounds of the clipping passes (sourcefind.cpp:39-151) with numpy sums: only
    used to show which passes of a case read the candidates, with a wide margin."""
    v = img.ravel(order="F") - np.median(img)
    keep = np.ones(v.size, bool)
    bounds = [num_sigma * img.std()]
    for _ in range(iters):
        rem = keep & (np.abs(v) > bounds[-1])
        if not rem.any():
            break
        keep &= ~rem
        bounds.append(num_sigma * v[keep].std())
    return bounds


@pytest.mark.parametrize("case", ["noise", "bright", "ragged"])
def test_gpu_sigma_clip_candidates(gpu, orc, case):
    """Clipping passes after a full one read only the pixels beyond 0.5 x the full
    pass's bound; once the bound falls below that, a pass runs over the whole image
    again.  noise: every later pass reads candidates; bright: a few very bright
    pixels make the first bound ~10x the next (a second full pass, then candidates);
    ragged: a partial last canonical block (and sources that move the bound by 15 %).  Bar: statistics and catalogue bit-exact."""
    rng = np.random.default_rng(21)
    if case == "ragged":
        img = _field(300, 302, seed=21)
    else:
        img = np.asfortranarray(noise(512, 512, sigma=1.0, seed=21))
        if case == "bright":
            img.ravel(order="F")[rng.choice(img.size, 60, replace=False)] = 500.0
    iters = 12
    b = _clip_bounds(img, iters)
    assert len(b) >= 3
    if case == "bright":
        assert b[1] < 0.5 * 0.5 * b[0] and all(b[k + 1] > 0.97 * b[k] for k in range(1, len(b) - 1))
    elif case == "noise":
        assert all(b[k] > 0.95 * b[0] for k in range(len(b)))
    else:
        assert all(b[k] > 0.6 * b[0] for k in range(len(b)))
    pars = capi.sourcefind_pars(5.0, 3.0, sigma_clip_iters=iters, median_method=capi.MEDIAN_BINMEDIAN)
    compare(gpu.source_find(img, pars), orc.source_find(img, pars))


def test_gpu_island_list_pinned_buffer(gpu, orc):
    """An island list written by DMA into a page-locked buffer (copied with the state
    while the label map runs) equals the one through ordinary memory and the
    oracle's: small, large, small catalogues on one context, so that the copy sized
    by the previous call's count falls short, then overshoots."""
    fields = [_field(256, 256, seed=41, n_src=2), _field(512, 512, seed=42, n_src=60), _field(256, 256, seed=43, n_src=3)]
    pars = capi.sourcefind_pars(4.0, 2.5, median_method=capi.MEDIAN_BINMEDIAN, source_min_area=1)
    pinned = gpu.island_buffer(1 << 14)
    try:
        counts = []
        for img in fields:
            rows, cols = img.shape
            d_img = gpu.malloc(8 * rows * cols)
            try:
                gpu.htod(d_img, img)
                a, sa = gpu.source_find_device(d_img, rows, cols, pars, precision=capi.PRECISION_FP64, out=pinned)
                a = list(a)
                b, sb = gpu.source_find_device(d_img, rows, cols, pars, precision=capi.PRECISION_FP64, capacity=1 << 14)
                b = list(b)
            finally:
                gpu.free(d_img)
            assert sa.num_islands == sb.num_islands == len(a) == len(b)
            _, oisl, _ = orc.source_find(img, pars)
            assert len(a) == len(oisl)
            # moments: sums in atomic order, so within fits_agree() from call to call
            for x, y, z in zip(a, b, oisl):
                assert all(x[f] == y[f] == z[f] for f in ISLAND_EXACT)
                assert fits_agree(x["moments_fit"], y["moments_fit"]) and fits_agree(x["moments_fit"], z["moments_fit"])
            counts.append(len(a))
        assert counts[1] > 2 * counts[0] + 64 and counts[2] < counts[1]
    finally:
        gpu.free_host(pinned)


def test_gpu_island_capacity_too_small(gpu):
    """A catalogue larger than the island buffer: STP_ERR_CAPACITY naming the count, and the
    first `capacity` islands delivered, into a page-locked buffer (written by the emitting
    kernel, which stops at the capacity) as into an ordinary one."""
    from fastimaging_amd.stp import STPError
    img = _field(512, 512, seed=42, n_src=60)
    rows, cols = img.shape
    pars = capi.sourcefind_pars(4.0, 2.5, median_method=capi.MEDIAN_BINMEDIAN, source_min_area=1)
    d_img = gpu.malloc(8 * rows * cols)
    pinned = gpu.island_buffer(32)
    try:
        gpu.htod(d_img, img)
        full, st = gpu.source_find_device(d_img, rows, cols, pars, precision=capi.PRECISION_FP64, capacity=1 << 12)
        full = list(full)
        assert len(full) > 64
        plain = (capi.Island * 32)()
        for buf in (pinned, plain):
            with pytest.raises(STPError) as e:
                gpu.source_find_device(d_img, rows, cols, pars, precision=capi.PRECISION_FP64, out=buf)
            assert e.value.status == 6 and f"{len(full)} islands found" in str(e.value)
            got = [capi.island_tuple(buf[i]) for i in range(32)]
            for x, y in zip(got, full[:32]):
                assert all(x[f] == y[f] for f in ISLAND_EXACT) and fits_agree(x["moments_fit"], y["moments_fit"])
    finally:
        gpu.free(d_img)
        gpu.free_host(pinned)


def test_gpu_sourcefind_degenerate(gpu, orc):
    """Constant image (sigma 0: exact-median fallback), sparse image (most pixels exactly 0:
    radix selection), all-source image (every pixel above threshold)."""
    pars = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN, source_min_area=1)
    c = np.full((64, 64), 2.5, order="F")
    compare(gpu.source_find(c, pars), orc.source_find(c, pars))
    s = np.zeros((128, 128), order="F")
    rng = np.random.default_rng(5)
    idx = rng.choice(s.size, 700, replace=False)
    s.ravel(order="F")[idx] = rng.normal(size=700) * 10
    s = np.asfortranarray(s)
    for m in (capi.MEDIAN_BINMEDIAN, capi.MEDIAN_BINAPPROX, capi.MEDIAN_NTHELEMENT):
        pars.median_method = m
        compare(gpu.source_find(s, pars), orc.source_find(s, pars))
    pars = capi.sourcefind_pars(1.0, 0.5, rms_est=1.0, median_method=capi.MEDIAN_ZEROMEDIAN, source_min_area=1)
    a = np.full((32, 32), 10.0, order="F")
    a[5, 5] = 11.0
    compare(gpu.source_find(a, pars), orc.source_find(a, pars))


@pytest.mark.parametrize("conn4", [False, True])
def test_gpu_sourcefind_dense_components(gpu, orc, conn4):
    """Noise labelled at 1 sigma: 500-600 components per 64 x 64 tile, more than a tile
    sums in LDS, so part of every tile's islands take the global-atomic path; and
    islands spread over many tiles (a 1-sigma threshold percolates)."""
    img = noise(256, 320, seed=21)
    pars = capi.sourcefind_pars(1.5, 1.0, rms_est=1.0, median_method=capi.MEDIAN_ZEROMEDIAN, ccl_4connectivity=conn4)
    compare(gpu.source_find(img, pars, capacity=1 << 15), orc.source_find(img, pars, capacity=1 << 15))


def test_gpu_sourcefind_component_records_grow(gpu, orc):
    """A +-2 checkerboard with 4-connectivity makes every pixel its own island: 262144
    components in a 512 x 512 image, more than the initial component-record capacity,
    so the labelling grows it and runs again."""
    cb = np.asfortranarray(2.0 * ((np.indices((512, 512)).sum(axis=0) % 2) * 2 - 1))
    pars = capi.sourcefind_pars(1.5, 1.0, rms_est=1.0, median_method=capi.MEDIAN_ZEROMEDIAN, ccl_4connectivity=True)
    for _ in range(2):  # the grown capacity is kept by the next call
        res = gpu.source_find(cb, pars)
        compare(res, orc.source_find(cb, pars))
        assert res.stats.num_labels_pos == res.stats.num_labels_neg == 512 * 512 // 2


@pytest.mark.parametrize("median", [capi.MEDIAN_BINMEDIAN, capi.MEDIAN_NTHELEMENT])
def test_gpu_exact_median_two_level(gpu, orc, median):
    """A peaked distribution (99% of pixels with sigma 1e-3, 1% with sigma 100) puts
    ~all pixels in the first-level bin of the median, so the selection refines
    through the second level; the median must still be the exact order statistic."""
    rng = np.random.default_rng(42)
    n = 1024
    img = rng.normal(scale=1e-3, size=(n, n))
    wide = rng.random(size=(n, n)) < 0.01
    img[wide] = rng.normal(scale=100.0, size=int(wide.sum()))
    img = np.asfortranarray(img)
    pars = capi.sourcefind_pars(5.0, 3.0, median_method=median)
    res = gpu.source_find(img, pars)
    assert res.stats.bg_level == np.median(img)
    compare(res, orc.source_find(img, pars))


def test_gpu_source_detection_kats(gpu):
    """sourcefind_test_BasicSourceDetection / NegativeSourceDetection through the GPU."""
    for sign, neg in ((1.0, False), (-1.0, True)):
        ydim, xdim = 128, 64
        bright = gaussian2d(ydim, xdim, 48.24, 52.66, sign * 10.0, 1.5, 1.2, 1.0)
        faint = gaussian2d(ydim, xdim, 32, 64, sign * 3.5, 1.5, 1.2, 1.0)
        p = capi.sourcefind_pars(4, 3, rms_est=1.0, find_negative_sources=neg)
        r0 = gpu.source_find(fftshift(bright), p)
        r1 = gpu.source_find(fftshift(bright + faint), p)
        r2 = gpu.source_find(fftshift(bright + 2 * faint), p)
        assert len(r0.islands) == 1 and len(r1.islands) == 1 and len(r2.islands) == 2
        assert abs(r0.islands[0]["extremum_x_idx"] - 48.24) < 0.5
        assert abs(r0.islands[0]["extremum_y_idx"] - 52.66) < 0.5


def test_gpu_sourcefind_many_canonical_blocks(gpu, orc):
    """6144 x 6144 pixels, more than 4096 canonical blocks: the statistics and the clipping
    passes reduce their block partials through an extra canonical level.  Noise with 40 very
    bright 3 x 3 blocks, so the first clipping bound is far above the next (a second full
    pass, then candidate passes).  Bar: compare() (statistics bit-exact)."""
    img = np.asfortranarray(noise(6144, 6144, sigma=1.0, seed=5))
    rng = np.random.default_rng(5)
    for _ in range(40):
        r, c = rng.integers(8, 6136, 2)
        img[r - 1:r + 2, c - 1:c + 2] += rng.uniform(2000.0, 8000.0)
    b = _clip_bounds(img, 5)
    assert b[1] < 0.25 * b[0]
    pars = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    compare(gpu.source_find(img, pars), orc.source_find(img, pars))


def test_gpu_pipeline_c2_full_size(gpu, orc):
    """Workload c2 at full size: the oracle image through both source finders is bit-exact,
    and the GPU image gives the same islands as the oracle image."""
    snap = simdata.workload_snapshot("c2")
    pars = capi.imager_pars(4096, 1.0)
    img, _ = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars)
    oimg, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    ores = orc.source_find(oimg, sfp)
    compare(gpu.source_find(oimg, sfp), ores)
    g2 = gpu.source_find(img, sfp)
    assert [(i["label_idx"], i["extremum_x_idx"], i["extremum_y_idx"]) for i in g2.islands] == \
           [(i["label_idx"], i["extremum_x_idx"], i["extremum_y_idx"]) for i in ores[1]]
    assert len(ores[1]) >= 40
