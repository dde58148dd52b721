"""Pins of the oracle's Hankel W-kernels (aw_projection.cpp:211-287,380-411,477-596;
aw_projection.h:127-213; the vendored tk::spline, src/stp/common/spline.h).  The
reference has no Hankel test, so each step is checked against what fixes it
independently: J1 against scipy; the DHT of a Gaussian against its closed-form
Hankel transform; the natural cubic and linear splines against scipy and numpy;
the projection-slice profile against the 1-D DFT of the 2-D image-domain kernel
summed over one axis (the projection-slice theorem); and every cache element
against the spline of the returned profile at radius sqrt(cx^2 + cy^2), normalised
by its sub-kernel's real sum, plus the truncation rule."""
import numpy as np
import pytest
import scipy.interpolate
import scipy.special


def test_besselj1_matches_scipy(orc):
    x = np.linspace(0.0, 200.0, 40001)
    assert np.max(np.abs(orc.besselj1(x) - scipy.special.j1(x))) < 2e-8


def test_spline_natural_cubic_and_linear(orc):
    rng = np.random.default_rng(5)
    for n in (3, 4, 17, 64):
        y = rng.normal(size=n)
        xs = rng.uniform(0, n - 1, 400)
        cs = scipy.interpolate.CubicSpline(np.arange(n), y, bc_type="natural")
        assert np.max(np.abs(orc.spline(y, xs, True) - cs(xs))) < 1e-12
        assert np.max(np.abs(orc.spline(y, xs, False) - np.interp(xs, np.arange(n), y))) < 1e-14
        # beyond the last knot: the line with the last knot's slope
        xe = np.array([n - 0.5, n + 2.0])
        assert np.allclose(orc.spline(y, xe, True), y[-1] + cs(n - 1, 1) * (xe - (n - 1)), rtol=0, atol=1e-12)
        assert np.allclose(orc.spline(y, xe, False), y[-1] + (y[-1] - y[-2]) * (xe - (n - 1)), rtol=0, atol=1e-13)


def test_dht_of_gaussian(orc):
    """Column i of the DHT integrates 2 pi r J0(k_i r) over ring j (k_i = pi i / n): the
    transform of exp(-r^2 / 2 s^2) is 2 pi s^2 exp(-s^2 k^2 / 2)."""
    n, s = 256, 8.0
    D = orc.dht_matrix(n)
    f = np.exp(-np.arange(n) ** 2 / (2 * s * s))
    k = np.pi * np.arange(n) / n
    F = 2 * np.pi * s * s * np.exp(-s * s * k * k / 2)
    assert np.max(np.abs(D.T @ f - F)) < 1e-3 * F[0]
    # column 0: the ring areas, pi rn_j^2 - pi rn_{j-1}^2 (rn_j = j + 1/2, the last n - 1)
    rn = np.arange(n) + 0.5
    rn[-1] = n - 1
    assert np.allclose(D[:, 0], np.diff(np.concatenate([[0.0], np.pi * rn * rn])), rtol=1e-14, atol=0)


def _setup(orc, workarea=64, os_=4, N=256, cell_arcsec=20.0):
    aa = orc.img_dom_kernel(4, 3.0, workarea, widths=(2.52, 1.55))  # GaussianSinc image-domain kernel
    cell_rad = cell_arcsec / 3600.0 * np.pi / 180.0
    return aa, cell_rad, N / workarea


def _expected_cache(rad, ts, os_, interp):
    """Cache from the profile: element (row tx, col ty) of sub-kernel (x, y) is the profile
    at radius |(cx, cy)|, cx = -ts os + os/2 - x + ty os, cy likewise with y and tx."""
    op = os_ // 2
    cs = 2 * op + 1
    kt = 2 * ts + 1
    out = np.empty((cs, cs, kt, kt), dtype=np.complex128)
    for x in range(cs):
        for y in range(cs):
            t = np.arange(kt)
            cx = -ts * os_ + op - x + t * os_
            cy = -ts * os_ + op - y + t * os_
            r = np.sqrt(cy[:, None] ** 2 + cx[None, :] ** 2)
            v = interp(r)
            out[y, x] = v / v.real.sum()
    return out


@pytest.mark.parametrize("proj_slice", [False, True])
@pytest.mark.parametrize("cubic", [False, True])
def test_hankel_cache_from_profile(orc, proj_slice, cubic):
    aa, cell_rad, scaling = _setup(orc)
    workarea, os_, support = 64, 4, 6
    cache, ts, rad = orc.hankel_kernel_cache(250.0, aa, workarea, cell_rad, os_, scaling, support, proj_slice=proj_slice,
                                             cubic=cubic)
    assert ts == support and rad.size == (workarea * os_ if proj_slice else workarea * os_ // 2)
    knots = np.arange(rad.size)
    if cubic:
        sr = scipy.interpolate.CubicSpline(knots, rad.real, bc_type="natural")
        si = scipy.interpolate.CubicSpline(knots, rad.imag, bc_type="natural")
        interp = lambda r: sr(r) + 1j * si(r)
    else:
        interp = lambda r: np.interp(r, knots, rad.real) + 1j * np.interp(r, knots, rad.imag)
    exp = _expected_cache(rad, ts, os_, interp)
    assert np.max(np.abs(cache - exp)) < 1e-11 * np.max(np.abs(exp))
    # every sub-kernel has unit real sum
    assert np.allclose(cache.real.sum(axis=(2, 3)), 1.0, atol=1e-12)


def test_projection_slice_profile(orc):
    """The projection-slice profile is the 1-D forward DFT of the 2-D image-domain kernel
    (screen e^{-2 pi i w (n - 1)} / n times aa(x) aa(y), mirrored into the four
    quadrants of the zero-padded array) summed over its second axis."""
    aa, cell_rad, scaling = _setup(orc)
    workarea, os_, support, w = 64, 4, 6, 300.0
    _, _, rad = orc.hankel_kernel_cache(w, aa, workarea, cell_rad, os_, scaling, support, proj_slice=True)
    arr, h = workarea * os_, workarea // 2
    sc = cell_rad * scaling
    comb = np.zeros((arr, arr), dtype=np.complex128)
    for i in range(h):
        for j in range(h):
            r2 = (i * sc) ** 2 + (j * sc) ** 2
            v = aa[h + i] * aa[h + j]
            if r2 < 1.0:
                n = np.sqrt(1.0 - r2)
                v = v * np.exp(-2j * np.pi * w * (n - 1.0)) / n
            for a in {i, (arr - i) % arr}:
                for b in {j, (arr - j) % arr}:
                    comb[a, b] = v
    want = np.fft.fft(comb.sum(axis=1))
    assert np.max(np.abs(rad - want)) < 1e-10 * np.max(np.abs(want))


def test_hankel_truncation_rule(orc):
    """trunc_perc > 0: the largest support s (>= 1) whose profile sample at s os exceeds
    trunc_perc % of the sample at 0, searched down from the maximum support."""
    aa, cell_rad, scaling = _setup(orc)
    workarea, os_, support = 64, 4, 14
    for perc in (0.5, 5.0, 30.0):
        _, ts, rad = orc.hankel_kernel_cache(400.0, aa, workarea, cell_rad, os_, scaling, support, trunc_perc=perc)
        idx = support * os_
        while idx > os_ and not abs(rad[idx]) > abs(rad[0]) * perc / 100.0:
            idx -= os_
        assert ts == min(int(np.ceil(idx / os_)), support) and ts >= 1
