"""Pins of the oracle's W-projection restatement (gridder.cpp:68-146,
aw_projection.cpp:19-210,366-474, gridder.h:311-367,767-842).  The reference has no
W-projection test, so these check what the mathematics fixes: equal-count planes and
their mean/median |w|; the W-kernel for w = 0 is the separable product of one 1-D
FFT of the image-domain kernel and tracks the oversampled kernel cache; w -> -w
conjugates the kernel; kernels of equal x/y offsets are symmetric; every kernel has
unit real sum; and W-projection refocuses a far-field point source that plain
gridding smears, at the source's own pixel."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import peak_rel_err


def test_average_w_planes_equal_counts(orc):
    rng = np.random.default_rng(3)
    n, P = 1000, 7
    w = np.sort(rng.uniform(0, 1500, n))
    avg, first = orc.average_w_planes(w, np.ones(n, np.uint32), P)
    bounds = list(first.astype(int)) + [n]
    sizes = np.diff(bounds)
    # ceil((n - begin) / (P - p)): the first planes take the larger share, sizes differ by <= 1
    assert sizes.sum() == n and sizes.max() - sizes.min() <= 1 and np.all(np.diff(sizes) <= 0)
    for p in range(P):
        assert abs(avg[p] - w[bounds[p]:bounds[p + 1]].mean()) <= 1e-12 * w.max()
    med, _ = orc.average_w_planes(w, np.ones(n, np.uint32), P, median=True)
    for p in range(P):
        seg = w[bounds[p]:bounds[p + 1]]
        assert med[p] == pytest.approx(np.median(seg), rel=0, abs=1e-12 * w.max())


def test_average_w_planes_skips_bad(orc):
    """Bad visibilities (good = 0) count for no plane; placed after every good one they
    leave the planes of the good visibilities unchanged."""
    rng = np.random.default_rng(4)
    n_good, n_bad, P = 600, 40, 5
    w = np.sort(rng.uniform(0, 100, n_good + n_bad))
    good = np.ones(n_good + n_bad, np.uint32)
    good[n_good:] = 0
    avg, first = orc.average_w_planes(w, good, P)
    avg0, first0 = orc.average_w_planes(w[:n_good], good[:n_good], P)
    assert np.array_equal(first[:P], first0) and np.allclose(avg, avg0, rtol=0, atol=1e-12)


def _aa(orc, workarea):
    return orc.img_dom_kernel(capi.KERNEL_GAUSSIANSINC, 3.0, workarea, normalize=False, analytic=False)


@pytest.mark.parametrize("support,os_", [(8, 8), (6, 4)])
def test_wkernel_w0_separable(orc, support, os_):
    """w = 0: the image-domain kernel is aa(x) aa(y) on the +-(workarea/2 - 1) corners of
    the oversampled array (aw_projection.cpp:79-108), so its 2-D FFT is the outer
    product of one 1-D FFT (numpy here) and every cut kernel is that product sampled
    at (tap - s) * os - (offset - os/2), normalised to unit sum.  The cut kernels also
    track the oversampled kernel cache of the same functor (the interpolant of the
    integer samples against the function itself), which fixes the direction of the
    sub-pixel offsets."""
    ks = 2 * support + 1
    workarea = 1
    while workarea < ks:
        workarea *= 2
    aa = _aa(orc, workarea)
    # a vanishing cell keeps the 1/n factor of the image-domain kernel at 1
    cache = orc.wproj_kernel_cache(0.0, aa, workarea, 1e-12, os_, 1.0, support)
    arr, half, op = workarea * os_, workarea // 2, os_ // 2
    a = np.zeros(arr)
    a[0] = aa[half]
    for p in range(1, half):
        a[p] = a[arr - p] = aa[half + p]
    A = np.fft.fft(a)
    ref = orc.populate_kernel_cache(capi.KERNEL_GAUSSIANSINC, 3.0, support, os_)
    taps = np.arange(ks)
    for y in range(os_ + 1):
        for x in range(os_ + 1):
            cx = A[((taps - support) * os_ + op - x) % arr]
            cy = A[((taps - support) * os_ + op - y) % arr]
            expect = np.outer(cy, cx)
            expect = expect / expect.real.sum()
            got = cache[y, x]
            assert np.max(np.abs(got - expect)) < 1e-12, (x, y)
            b = ref[y, x].real
            assert np.max(np.abs(got.real - b)) < 2e-2 * np.max(np.abs(b)), (x, y)


def test_wkernel_conjugate_symmetry_and_sum(orc):
    support, os_, workarea = 8, 8, 32
    aa = _aa(orc, workarea)
    cell_rad = 20.0 / 3600 * np.pi / 180
    scaling = 256.0 / workarea
    c = orc.wproj_kernel_cache(250.0, aa, workarea, cell_rad, os_, scaling, support)
    cn = orc.wproj_kernel_cache(-250.0, aa, workarea, cell_rad, os_, scaling, support)
    assert np.max(np.abs(c - np.conj(cn))) < 1e-14
    assert np.max(np.abs(c.imag)) > 1e-3  # the w-term makes the kernel complex
    for y in range(os_ + 1):
        for x in range(os_ + 1):
            assert abs(c[y, x].real.sum() - 1.0) < 1e-12
        assert np.max(np.abs(c[y, y] - c[y, y].T)) < 1e-14


@pytest.mark.parametrize("w_value,trunc_perc", [(250.0, 1.0), (900.0, 1.0), (50.0, 5.0), (250.0, 0.01)])
def test_wkernel_truncation(orc, w_value, trunc_perc):
    """kernel_trunc_perc (aw_projection.cpp:176-179,197-209, :366-452): the plane's support
    is the largest s whose tap (row S, column S + s) of the centred kernel exceeds
    trunc_perc % of its centre tap, in magnitude (that tap is the FFT sample conv(s os, 0)
    the reference tests, scaled by the kernel's real sum), and each truncated kernel is the
    centred window of the full one renormalised to a unit real sum."""
    support, os_, workarea = 12, 8, 32
    aa = _aa(orc, workarea)
    cell_rad = 20.0 / 3600 * np.pi / 180
    scaling = 256.0 / workarea
    full = orc.wproj_kernel_cache(w_value, aa, workarea, cell_rad, os_, scaling, support)
    trunc, ts = orc.wproj_kernel_cache(w_value, aa, workarea, cell_rad, os_, scaling, support, trunc_perc=trunc_perc,
                                       return_support=True)
    op, S = os_ // 2, support
    centre = full[op, op]
    mags = np.abs(centre[S, S:])
    expect = max([s for s in range(1, S + 1) if mags[s] > trunc_perc / 100.0 * mags[0]] + [0])
    assert ts == expect and 0 < ts <= S
    if trunc_perc >= 1.0:
        assert ts < S  # the test cases do truncate
    assert trunc.shape[2:] == (2 * ts + 1, 2 * ts + 1)
    for y in range(os_ + 1):
        for x in range(os_ + 1):
            win = full[y, x][S - ts:S + ts + 1, S - ts:S + ts + 1]
            want = win / win.real.sum()
            assert np.max(np.abs(trunc[y, x] - want)) < 1e-12 * np.max(np.abs(want)), (y, x)


def test_wprojection_refocuses_point_source(orc):
    """A 10 Jy source 1.4 deg from the phase centre, |w| up to 1500, visibilities from the
    reference's own point-source model (visibility.cpp:15-28): plain gridding smears it,
    W-projection restores the peak at the source's pixel (-l/cell, -m/cell in that
    sign convention) -- which fixes the sign of the W-kernel phase."""
    N, cell = 256, 120.0
    snap = simdata.snapshot(N, cell, 20000, 1, seed=42, amp_range=(10.0, 10.0), field_frac=0.2, w_max=1500.0,
                            noise_sigma=0.0, reference_model=True)
    A, l, m = snap["sources"][0]
    pars = capi.imager_pars(N, cell)
    plain, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
    wimg, _, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"],
                                           wproj=capi.wproj_pars(16, 16))
    cell_rad = cell / 3600 * np.pi / 180
    px, py = int(round(-l / cell_rad)) % N, int(round(-m / cell_rad)) % N
    iy, ix = np.unravel_index(np.argmax(wimg), wimg.shape)
    assert (iy, ix) == (py, px)
    assert wimg.max() > 3.0 * plain.max()
    assert wimg.max() > 0.7 * A
