"""Test helpers: fixtures in the style of src/auxiliary/fixtures.h and gaussian2d.cpp."""
import numpy as np


def fftshift(a):
    """stp::fftshift (src/stp/common/fft.h:92-108) for even-sized matrices."""
    return np.fft.fftshift(a)


def gaussian2d(n_rows, n_cols, x_mean, y_mean, amplitude, x_stddev, y_stddev, theta):
    """Gaussian2D::operator() (src/auxiliary/gaussian2d.cpp:4-31); column index = x."""
    cost, sint, sin2t = np.cos(theta), np.sin(theta), np.sin(2.0 * theta)
    xstd2, ystd2 = x_stddev * x_stddev, y_stddev * y_stddev
    a = 0.5 * ((cost * cost / xstd2) + (sint * sint / ystd2))
    b = 0.5 * ((sin2t / xstd2) - (sin2t / ystd2))
    c = 0.5 * ((sint * sint / xstd2) + (cost * cost / ystd2))
    x = np.arange(n_cols)[None, :] - x_mean
    y = np.arange(n_rows)[:, None] - y_mean
    return np.asfortranarray(amplitude * np.exp(-(a * x * x + b * x * y + c * y * y)))


def noise(rows, cols, sigma=1.0, mean=0.0, seed=0):
    """Uncorrelated Gaussian noise background (fixtures.h:19-32), numpy RNG."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(sigma * rng.standard_normal((rows, cols)) + mean)


def peak_rel_err(a, b):
    """max |a - b| / max |b|."""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
