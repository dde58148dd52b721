"""Pins the oracle's FFTs and statistics: FFTW c2r/c2c semantics against numpy.fft
and brute-force DFTs, medians against sorting (matrixmath_test_median_funcs.cpp),
mean/sigma against closed forms, sigma-clipped RMS (sourcefind_test_RmsEstimation.cpp)."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests.helpers import noise


def _half_grid(rng, n):
    return rng.normal(size=(n // 2 + 1, n)) + 1j * rng.normal(size=(n // 2 + 1, n))


@pytest.mark.parametrize("n", [8, 16, 64, 256])
def test_c2r_matches_numpy_irfftn(orc, n):
    """fft_fftw_c2r (fft.cpp:108-112): c2c along u, then c2r along v, unnormalised;
    numpy.fft.irfftn(G.T)*n^2 computes the same transform."""
    g = _half_grid(np.random.default_rng(n), n)
    out = orc.fft_c2r(g)
    ref = (np.fft.irfftn(g.T) * n * n).T
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref))


def test_c2r_brute_force_dft(orc):
    """Brute-force DFT with the Hermitian extension along v and Re of the v=0, n/2 rows."""
    n = 8
    g = _half_grid(np.random.default_rng(5), n)
    h = np.zeros((n, n), complex)
    for v in range(n // 2 + 1):
        h[v] = g[v]
    h[0] = h[0]
    out = np.zeros((n, n))
    # c2c along u first
    hu = np.array([[sum(g[v, u] * np.exp(2j * np.pi * u * x / n) for u in range(n)) for x in range(n)]
                   for v in range(n // 2 + 1)])
    for x in range(n):
        col = np.zeros(n, complex)
        col[0] = hu[0, x].real
        col[n // 2] = hu[n // 2, x].real
        for v in range(1, n // 2):
            col[v] = hu[v, x]
            col[n - v] = np.conj(hu[v, x])
        for y in range(n):
            out[y, x] = sum(col[v] * np.exp(2j * np.pi * v * y / n) for v in range(n)).real
    assert np.allclose(orc.fft_c2r(g), out, atol=1e-11)


def test_c2r_single_visibility_fringe(orc):
    """One grid value at (v0, u0) gives the cosine fringe 2 Re(G e^{2 pi i (u0 x + v0 y)/n})."""
    n, u0, v0 = 32, 5, 3
    g = np.zeros((n // 2 + 1, n), complex)
    g[v0, u0] = 0.7 - 0.2j
    out = orc.fft_c2r(g)
    y, x = np.mgrid[0:n, 0:n]
    ref = 2 * np.real((0.7 - 0.2j) * np.exp(2j * np.pi * (u0 * x + v0 * y) / n))
    assert np.allclose(out, ref, atol=1e-12)


def test_c2c_forward_matches_numpy(orc):
    a = np.random.default_rng(2).normal(size=(32, 32)) + 1j * np.random.default_rng(3).normal(size=(32, 32))
    assert np.allclose(orc.fft_c2c(a, True), np.fft.fft2(a), atol=1e-11)
    assert np.allclose(orc.fft_c2c(a, False), np.fft.ifft2(a) * a.size, atol=1e-11)


def test_img_dom_kernel_closed_form(orc):
    """ImgDomKernel DFT path (conv_func.h:361-402): aa[K/2+i] = sum_x k(x) cos(2 pi x i / K)
    over x in [-K/2, K/2) with k normalised to unit sum."""
    K, s = 64, 3
    aa = orc.img_dom_kernel(capi.KERNEL_GAUSSIANSINC, s, K, normalize=False, analytic=False)
    x = np.arange(K) - K // 2
    k = orc.kernel_eval(capi.KERNEL_GAUSSIANSINC, s, x.astype(float))
    k = k / k.sum()
    ref = np.empty(K)
    for j in range(K):
        i = j - K // 2
        ref[j] = np.sum(k * np.cos(2 * np.pi * x * i / K))
    assert np.allclose(aa, ref, atol=1e-13)


def test_canonical_sum_order_and_value(orc):
    """The canonical sum is within rounding of the exact sum (math.fsum) at any size."""
    import math
    rng = np.random.default_rng(11)
    for n in (1, 7, 4096, 4097, 300000):
        x = rng.normal(size=n) * 10 + 3
        s1, s2 = orc.canonical_pair_sum(x)
        assert abs(s1 - math.fsum(x)) <= 1e-12 * math.fsum(np.abs(x))
        assert abs(s2 - math.fsum(x * x)) <= 1e-12 * math.fsum(x * x)


def test_mean_stddev(orc):
    """mat_mean_and_stddev (matrix_math.cpp:457-499): population sigma."""
    x = noise(128, 96, sigma=2.0, mean=1.5, seed=4)
    m, s = orc.mean_stddev(x)
    assert abs(m - x.mean()) < 1e-13 and abs(s - x.std()) < 1e-12
    # matrixmath_test_stat_funcs: constant data has sigma 0
    m, s = orc.mean_stddev(np.full((10, 10), 3.25))
    assert m == 3.25 and s == 0.0


@pytest.mark.parametrize("seed", range(4))
def test_medians_against_sorting(orc, seed):
    """matrixmath_test_median_funcs.cpp:13-198: exact and bin medians equal the sort median;
    binapprox is within sigma/800."""
    x = noise(512, 512, sigma=1.0, seed=seed)
    ref = np.median(x)
    assert orc.median_exact(x) == ref
    assert orc.binmedian(x)[2] == ref
    assert abs(orc.binapprox(x)[2] - ref) < 1.0 / 800
    odd = x[:, :511]
    assert orc.median_exact(odd) == np.median(odd)


@pytest.mark.parametrize("pattern", [
    [[1, 30, 60]] * 3, [[1, 1, 60], [1, 1, 60], [1, 60, 60]], [[1, 1, 1], [1, 1, 1], [1, 1, 60]]])
def test_median_specific_patterns(orc, pattern):
    x = np.array(pattern, dtype=float)
    assert orc.binmedian(x)[2] == np.median(x)


def test_binapprox_degenerate_patterns(orc):
    """TestMedianSpecificPattern1/2 (matrix_math tests :29-52) and sigma ~ 0 fallback."""
    d = np.full((20, 20), 100.0)
    d[:, :10] = -100.0
    assert abs(orc.binapprox(d)[2] - np.median(d)) < 100.0 / 800
    d[0, 19] = 200.0
    d[19, 0] = -200.0
    assert abs(orc.binapprox(d)[2] - np.median(d)) < 100.0 / 800
    c = np.full((10, 10), 5.0)
    assert orc.binapprox(c)[2] == 5.0


def test_binapprox_bin_rule_brute_force(orc):
    """Median from the 1001-bin histogram with truncation toward zero (matrix_math.cpp:354-448),
    recomputed here for an odd-sized sample."""
    x = noise(101, 101, seed=9).ravel(order="F")
    m, s = orc.mean_stddev(x)
    scale = 1000.0 / (2 * s)
    left = m - s
    bins = np.zeros(1001, int)
    bottom = 0
    for v in x:
        b = int((v - left) * scale)
        if b < 0:
            bottom += 1
        elif b <= 1000:
            bins[b] += 1
    k = (x.size + 1) // 2
    cum = bottom + np.cumsum(bins)
    i = int(np.argmax(cum >= k))
    assert orc.binapprox(x)[2] == (i + 0.5) / scale + left


def test_rms_estimation_noise(orc):
    """sourcefind_test_RmsEstimation.cpp:43-60: RMS of N(0,1) noise within 5 %."""
    x = noise(256, 256, seed=1)
    med = orc.median_exact(x)
    m, s = orc.mean_stddev(x)
    rms = orc.estimate_rms(x, 3.0, 5, (m, s, med, 7))
    assert abs(rms - 1.0) < 0.05


def test_rms_clipping_against_direct_iteration(orc):
    """estimate_rms equals the explicit astropy-style clip: repeatedly drop |x-med| > 3 sigma."""
    x = noise(200, 200, seed=2)
    x[10:20, 10:20] += 40.0  # bright outliers
    xs = x.ravel()
    med = float(np.median(xs))
    keep = np.ones(xs.size, bool)
    sigma = float(np.sqrt(np.mean(xs * xs) - np.mean(xs) ** 2))
    for _ in range(5):
        new = keep & (np.abs(xs - med) <= 3 * sigma)
        if new.sum() == keep.sum():
            break
        keep = new
        d = xs[keep] - med
        sigma = float(np.sqrt(np.mean(d * d) - np.mean(d) ** 2))
    m, s = orc.mean_stddev(x)
    rms = orc.estimate_rms(x, 3.0, 5, (m, s, med, 7))
    assert abs(rms - sigma) < 1e-9
