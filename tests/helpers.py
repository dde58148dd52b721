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


ISLAND_EXACT = ("label_idx", "sign", "extremum_y_idx", "extremum_x_idx", "num_samples", "extremum_val")


def fits_agree(a, b, tol=1e-9, axes_tol=1e-6):
    """Moments fit (sourcefind.cpp:691-717).  Amplitude and barycentre within tol.
    The axes come from xx_bar - x_bar^2 (sourcefind.cpp:597-599), which cancels sums of
    order (N/2)^2 * flux, so the per-label summation order (unspecified in the reference,
    TBB) shows up there: squared axes are compared relative to semimajor^2 (a squared
    semiminor that rounds negative gives NaN on either side and counts as 0), and the
    angle only for clearly elongated islands, modulo pi."""
    a, b = np.array(a, dtype=float), np.array(b, dtype=float)
    if not np.allclose(a[:3], b[:3], rtol=tol, atol=tol):
        return False
    A2a, A2b = a[3] ** 2, b[3] ** 2
    B2a, B2b = np.nan_to_num(a[4] ** 2), np.nan_to_num(b[4] ** 2)
    if abs(A2a - A2b) > axes_tol * A2b or abs(B2a - B2b) > axes_tol * A2b:
        return False
    # line-shaped islands (semiminor ~ 0 or NaN): xy_bar is at rounding level and the
    # angle branch (sourcefind.cpp:707-711) is decided by its sign, on either side
    line = np.isnan(a[4]) or np.isnan(b[4]) or B2a <= 1e-6 * A2a or B2b <= 1e-6 * A2b
    if not line and (A2b - B2b) > 1e-3 * A2b:
        d = (a[5] - b[5] + np.pi / 2) % np.pi - np.pi / 2
        if abs(d) > axes_tol:
            return False
    return True


def compare(gpu_res, orc_res):
    lm, isl, st = orc_res
    assert np.array_equal(gpu_res.label_map, lm)
    g = gpu_res.stats
    for f in ("bg_level", "rms_est", "mean", "sigma"):
        a, b = getattr(g, f), getattr(st, f)
        assert a == b or (np.isnan(a) and np.isnan(b)), (f, a, b)  # bit-exact (NaN where the algebra gives NaN)
    for f in ("num_labels_pos", "num_labels_neg", "num_valid_labels", "num_islands"):
        assert getattr(g, f) == getattr(st, f), f
    assert len(gpu_res.islands) == len(isl)
    for a, b in zip(gpu_res.islands, isl):
        for f in ISLAND_EXACT:
            assert a[f] == b[f], f
        assert fits_agree(a["moments_fit"], b["moments_fit"]), (a, b)


def islands_agree_banded(isl, oisl, det_pos, det_neg, band):
    """Islands found on an FP32 image against the oracle's islands on the FP64 image.
    The two images differ by at most `band` per pixel, so an island whose extremum lies
    within `band` of its detection threshold (det_pos, or det_neg for negative islands)
    may be found on one side only.  Every island beyond the band must be on both sides
    with the same sign and extremum pixel, and its extremum value within `band`.
    Returns (islands beyond the band, islands inside it)."""
    def key(i):
        return (i["sign"], i["extremum_x_idx"], i["extremum_y_idx"])

    def clear(i):
        e = i["extremum_val"]
        return e > det_pos + band if i["sign"] > 0 else e < det_neg - band

    g = {key(i): i for i in isl}
    o = {key(i): i for i in oisl}
    strict = 0
    for mine, other in ((g, o), (o, g)):
        for k, i in mine.items():
            if clear(i):
                assert k in other, ("island beyond the guard band found on one side only", k, i["extremum_val"])
                assert abs(i["extremum_val"] - other[k]["extremum_val"]) <= band, k
                strict += 1
    banded = len(set(g) | set(o)) - len({k for k, i in g.items() if clear(i)} | {k for k, i in o.items() if clear(i)})
    return strict // 2, banded
