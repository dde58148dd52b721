"""Closed-form pins of two oracle pieces that the reference's own tests leave open.

1. The padded crop and gridding-correction division of normalise_image_beam_result_1D
   (src/stp/imager/imager.h:96-213, non-FFTSHIFT branch), with the image-domain kernel
   of conv_func.h:351-412.  With exact gridding and visibilities on integer grid points
   every visibility's kernel response is exactly the DFT gcf, so the corrected image of a
   point source is its dirty beam, (A/n) sum_k cos(2 pi (u_k dx + v_k dy) / K), at every
   output pixel, where (dx, dy) are signed pixel offsets from the source on the padded
   grid of size K.  The uncorrected image is that times gcf(x) gcf(y).
2. The moments fit of IslandParams::estimate_moments_fit (sourcefind.cpp:691-717) and the
   constrained conversion (fitting.cpp:32-57).  Pixels above the analysis threshold T of a
   Gaussian of peak A fill the ellipse x^2/a^2 + y^2/b^2 < 2 ln c, c = A / T.  Their
   flux-weighted second moment along an axis of standard deviation a is
   a^2 (1 - ln c / (c - 1)), so dividing by that factor recovers the axes; the angle is
   the direction of the major eigenvector of the Gaussian's covariance.
"""
import math

import numpy as np
import pytest

from fastimaging_amd import capi
from tests.helpers import fftshift, gaussian2d


def _signed(i, n):
    """Signed pixel offset from the phase centre of un-shifted row / column i of n."""
    return np.where(i < n // 2, i, i - n)


def _gcf_closed_form(orc, K):
    """aa[K/2 + i] = sum_x k(x) cos(2 pi x i / K) / sum_x k(x) over integer x (GaussianSinc,
    trunc 3, conv_func.cpp:92-103): the image-domain response of the sampled kernel."""
    x = np.arange(-3, 4, dtype=float)
    k = orc.kernel_eval(capi.KERNEL_GAUSSIANSINC, 3.0, x)
    i = np.arange(K) - K // 2
    return (k[None, :] * np.cos(2 * np.pi * np.outer(i, x) / K)).sum(axis=1) / k.sum()


def _exact_wavelengths(px, cell, K):
    """Wavelengths u with u * scale == px exactly in double (scale: the imager's
    wavelength-to-pixel factor, imager.h:293), or NaN where no double within 16 ulps has
    that product.  A fraction of even 1e-14 pixel moves the far tap of the exact kernel
    past its truncation radius (conv_func.cpp:92-103), which zeroes it."""
    scale = ((cell / 3600.0) * (np.pi / 180.0)) * K
    out = np.full(len(px), np.nan)
    for k, p in enumerate(px):
        lo = hi = p / scale
        for _ in range(16):
            if lo * scale == p:
                out[k] = lo
                break
            if hi * scale == p:
                out[k] = hi
                break
            lo, hi = np.nextafter(lo, -np.inf), np.nextafter(hi, np.inf)
    return out


def _integer_uv_point_source(N, K, cell, n, x0, y0, amp, seed):
    rng = np.random.default_rng(seed)
    r = 0.38 * K * np.sqrt(rng.random(n))
    ph = 2 * np.pi * rng.random(n)
    upx = np.rint(r * np.cos(ph))
    vpx = np.rint(r * np.sin(ph))
    u, v = _exact_wavelengths(upx, cell, K), _exact_wavelengths(vpx, cell, K)
    keep = ~(np.isnan(u) | np.isnan(v))
    assert keep.sum() > 0.3 * n
    upx, vpx = upx[keep], vpx[keep]
    uvw = np.asfortranarray(np.stack([u[keep], v[keep], np.zeros(len(upx))], axis=1))
    # source at signed padded-grid pixel (x0, y0): l = x0 * cell, m = y0 * cell
    vis = amp * np.exp(-2j * np.pi * (upx * x0 + vpx * y0) / K)
    return uvw, vis, upx, vpx


def _dirty_beam(N, K, upx, vpx, x0, y0, amp):
    dx = _signed(np.arange(N), N)[None, :] - x0  # columns: u / x
    dy = _signed(np.arange(N), N)[:, None] - y0  # rows: v / y
    acc = np.zeros((N, N))
    for u, v in zip(upx, vpx):
        acc += np.cos(2 * np.pi * (u * dx + v * dy) / K)
    return amp * acc / len(upx)


@pytest.mark.parametrize("N,pad", [(64, 2.0), (64, 1.0), (32, 4.0)])
def test_padded_crop_and_gcf_dirty_beam(orc, N, pad):
    K = int(N * pad)
    cell, amp, x0, y0 = 20.0, 3.5, -9, 13
    uvw, vis, upx, vpx = _integer_uv_point_source(N, K, cell, 800, x0, y0, amp, seed=N + K)
    w = np.ones(len(vis))
    expect = _dirty_beam(N, K, upx, vpx, x0, y0, amp)
    gcf = _gcf_closed_form(orc, K)
    g_row = gcf[K // 2 + _signed(np.arange(N), N)]
    common = dict(kernel_function=capi.KERNEL_GAUSSIANSINC, kernel_support=3, kernel_exact=True,
                  oversampling=1, generate_beam=True, analytic_gcf=False, padding_factor=pad)
    img_c, beam_c, _, ngood = orc.image_visibilities(capi.imager_pars(N, cell, gridding_correction=True, **common),
                                                    vis, w, uvw)
    assert ngood == len(vis)
    assert np.max(np.abs(img_c - expect)) < 1e-10 * amp
    # the source lands at un-shifted column x0 mod N, row y0 mod N with its full flux
    assert abs(img_c[y0 % N, x0 % N] - amp) < 1e-10 * amp
    assert np.max(np.abs(beam_c - _dirty_beam(N, K, upx, vpx, 0, 0, 1.0))) < 1e-10
    img_n, beam_n, _, _ = orc.image_visibilities(capi.imager_pars(N, cell, gridding_correction=False, **common),
                                                 vis, w, uvw)
    assert np.max(np.abs(img_n - expect * np.outer(g_row, g_row))) < 1e-10 * amp
    assert np.max(np.abs(beam_n - beam_c * np.outer(g_row, g_row))) < 1e-10


def test_padded_crop_matches_cropped_transform(orc):
    """The crop keeps the un-shifted layout: output pixel (i, j) is padded pixel
    (d(i) mod K, d(j) mod K) of the normalised padded transform, d the signed offset."""
    N, K, cell = 64, 256, 5.0
    rng = np.random.default_rng(3)
    n = 3000
    r = 0.4 / (cell * math.pi / 180.0 / 3600.0) * np.sqrt(rng.random(n))
    ph = 2 * np.pi * rng.random(n)
    uvw = np.asfortranarray(np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(n)], axis=1))
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w = rng.random(n) + 0.5
    pars = capi.imager_pars(N, cell, kernel_function=capi.KERNEL_GAUSSIANSINC, kernel_support=3,
                            kernel_exact=False, oversampling=9, generate_beam=False, padding_factor=K / N)
    img, _, total, _ = orc.image_visibilities(pars, vis, w, uvw)
    scale = ((cell / 3600.0) * (np.pi / 180.0)) * K
    uv = np.asfortranarray(uvw[:, :2] * scale)
    grid, _, total2, _ = orc.convolve_to_grid(capi.KERNEL_GAUSSIANSINC, 3.0, 3, K, uv, vis, w, kernel_exact=False,
                                              oversampling=9, generate_beam=False)
    assert total == total2
    padded = orc.fft_c2r(grid) * (1.0 / total2)
    idx = _signed(np.arange(N), N) % K
    assert np.array_equal(img, padded[np.ix_(idx, idx)])


def _true_axes(sx, sy, theta):
    """Covariance of gaussian2d(.., sx, sy, theta): exp(-(a x^2 + b x y + c y^2)) has
    covariance (2 Q)^-1 with Q = [[a, b/2], [b/2, c]]."""
    ct, st, s2 = math.cos(theta), math.sin(theta), math.sin(2 * theta)
    a = 0.5 * (ct * ct / sx ** 2 + st * st / sy ** 2)
    b = 0.5 * (s2 / sx ** 2 - s2 / sy ** 2)
    c = 0.5 * (st * st / sx ** 2 + ct * ct / sy ** 2)
    cov = np.linalg.inv(2.0 * np.array([[a, b / 2], [b / 2, c]]))
    ev, vec = np.linalg.eigh(cov)
    major = vec[:, 1]
    ang = math.atan2(major[1], major[0])  # x = column, y = row (visual coordinates)
    ang = (ang + math.pi / 2) % math.pi - math.pi / 2
    if ang <= -math.pi / 2:
        ang += math.pi
    return math.sqrt(ev[1]), math.sqrt(ev[0]), ang


@pytest.mark.parametrize("sx,sy,theta", [(8.0, 4.0, 0.3), (8.0, 4.0, 1.2), (8.0, 4.0, -0.4), (8.0, 4.0, -1.3),
                                         (4.0, 8.0, 0.3), (4.0, 7.0, -1.1), (7.0, 3.5, 1.55), (7.0, 3.5, -1.55),
                                         (6.0, 5.0, 0.0)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_moments_fit_recovers_gaussian(orc, sx, sy, theta, sign):
    amp, rms, ana = 30.0, 1.0, 3.0  # cutoff ratio c = 10: the truncation factor 1 - ln c / (c - 1) = 0.744
    img = gaussian2d(160, 144, 70.3, 81.6, sign * amp, sx, sy, theta)
    p = capi.sourcefind_pars(5, ana, rms_est=rms, find_negative_sources=True)
    _, isl, _ = orc.source_find(fftshift(img), p)
    assert len(isl) == 1
    g = isl[0]["moments_fit"]
    A, B, th = _true_axes(sx, sy, theta)
    assert g[0] == isl[0]["extremum_val"]
    assert abs(g[1] - 70.3) < 0.05 and abs(g[2] - 81.6) < 0.05  # edge pixelisation
    # pixelisation of the truncation edge costs < 1.5 %; a dropped or wrong truncation
    # correction is off by 14 %
    assert abs(g[3] / A - 1.0) < 0.015 and abs(g[4] / B - 1.0) < 0.015
    assert g[3] >= g[4]
    if abs(A - B) > 0.1 * A:
        d = (g[5] - th + math.pi / 2) % math.pi - math.pi / 2
        assert abs(d) < 0.01, (g[5], th)
    assert -math.pi / 2 < g[5] <= math.pi / 2


def test_moments_fit_axis_order_and_angle_fold(orc):
    """A Gaussian elongated along the rows (x_stddev < y_stddev, theta = 0) has its major
    axis at +-pi/2: the swap branch adds pi/2 and the fold maps it into (-pi/2, pi/2]."""
    img = gaussian2d(128, 128, 60.0, 64.0, 40.0, 3.0, 9.0, 0.0)
    p = capi.sourcefind_pars(5, 3, rms_est=1.0, find_negative_sources=False)
    _, isl, _ = orc.source_find(fftshift(img), p)
    g = isl[0]["moments_fit"]
    assert abs(abs(g[5]) - math.pi / 2) < 1e-6
    assert g[3] > 2.5 * g[4]
