"""GPU island fit (stp_fit_islands[_device]) against the oracle's fit and the
reference's fitting test (src/test/sourcefind/sourcefind_test_Fitting.cpp:44-123).

Tolerances: islands of at least 12 samples whose fits converge agree with the oracle's
to 1e-6 (amplitude relative, positions and axes in pixels, angle in radians, for
elongated islands) and in cost to 1e-6 (relative); islands of 8 to 11 samples agree in
cost to 1e-3.  The fit is non-convex: where the two runs, which sum in different orders,
end at different points, the GPU's must be a stationary point or no worse than the
oracle's, and such islands stay under 1 %.  Islands whose fit reaches the iteration
limit (the minimum lies at infinity along a valley: the axes grow without bound) and
islands of 5 to 7 samples (6 parameters: flat or degenerate minima) need only finite
results.  Bounding boxes and the results of islands too small to solve are exact."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import fftshift, gaussian2d

pytestmark = pytest.mark.gpu


def _pixels(img, lm, isl):
    """The island's samples in visual coordinates (x = column, y = row)."""
    rows, cols = img.shape
    r, c = np.nonzero(lm == isl["label_idx"])
    y = np.where(r < rows // 2, r + rows // 2, r - rows // 2).astype(float)
    x = np.where(c < cols // 2, c + cols // 2, c - cols // 2).astype(float)
    return x, y, img[r, c]


def _cost(img, lm, isl, p):
    """Sum of squared residuals of fit p over the island's pixels."""
    x, y, v = _pixels(img, lm, isl)
    A, x0, y0, sx, sy, t = p
    ct, st, s2t = np.cos(t), np.sin(t), np.sin(2 * t)
    a = 0.5 * (ct * ct / sx ** 2 + st * st / sy ** 2)
    b = 0.5 * (s2t / sx ** 2 - s2t / sy ** 2)
    cc = 0.5 * (st * st / sx ** 2 + ct * ct / sy ** 2)
    dx, dy = x - x0, y - y0
    return float(np.sum((A * np.exp(-(a * dx * dx + b * dx * dy + cc * dy * dy)) - v) ** 2))


def _stationary(orc, img, lm, isl, p):
    """Whether p is a stationary point of the island's sum of squares: J^T r = 0 relative
    to the size of its terms (the oracle's analytic Jacobian, pinned by finite differences)."""
    grad, scale = np.zeros(6), np.zeros(6)
    for x, y, v in zip(*_pixels(img, lm, isl)):
        r, J = orc.gaussian_jacobian(p, x, y, v)
        grad += J * r
        scale += np.abs(J * r)
    return bool(np.all(np.abs(grad) <= 1e-6 * scale + 1e-12))


def _compare(orc, img, lm, isl, got, want):
    """Both fits minimise the same sum of squares.  Islands of >= 8 samples: the costs
    agree (1e-6 relative from 12 samples, 1e-3 below); where they do not, the two runs
    reached different local minima of a non-convex problem, which is accepted when the
    GPU's is itself stationary or no worse than the oracle's, and such islands are
    under 1 %.  Islands of >= 12
    samples with agreeing costs: parameters within 1e-6."""
    assert len(got) == len(want) == len(isl)
    n_cmp = n_tight = n_unconverged = 0
    other_minimum, bad = [], []
    worst_cost = worst_par = 0.0
    for k, (i, g, w) in enumerate(zip(isl, got, want)):
        assert g["bounding_box"] == w["bounding_box"], k
        if w["status"] == 2:
            assert g["status"] == 2 and g["leastsq_fit"] == w["leastsq_fit"], k
            continue
        assert np.all(np.isfinite(g["leastsq_fit"])), k
        m = i["num_samples"]
        if g["status"] == 1 or w["status"] == 1:
            # the iteration limit: the minimum lies at infinity along a valley (axes
            # growing without bound), where each run stops at a different point
            n_unconverged += 1
            continue
        if m < 8:
            # 5 to 7 samples for 6 parameters: the minimum is a flat valley (or a
            # zero-residual family); where each run stops depends on rounding
            continue
        n_cmp += 1
        cg, cw = _cost(img, lm, i, g["leastsq_fit"]), _cost(img, lm, i, w["leastsq_fit"])
        rc = abs(cg - cw) / max(cw, 1e-300)
        if rc > (1e-6 if m >= 12 else 1e-3) and abs(cg - cw) > 1e-20:
            # different end points: accepted when the GPU's is a stationary point or no
            # worse than the oracle's
            if cg <= cw * (1 + 1e-9) or _stationary(orc, img, lm, i, np.array(g["leastsq_fit"])):
                other_minimum.append((k, m, cg, cw))
            else:
                bad.append(("worse and not a minimum", k, m, cg, cw))
            continue
        worst_cost = max(worst_cost, rc)
        if m >= 12 and w["status"] == 0 and g["status"] == 0:
            a, b = np.array(g["leastsq_fit"]), np.array(w["leastsq_fit"])
            d = np.abs(a - b) / np.r_[abs(b[0]), np.maximum(1.0, np.abs(b[1:5])), 1.0]
            if b[3] - b[4] <= 1e-3 * b[3]:
                d[5] = 0.0  # nearly circular: the angle is not determined
            else:
                d[5] = abs((a[5] - b[5] + np.pi / 2) % np.pi - np.pi / 2)
            worst_par = max(worst_par, float(d.max()))
            if d.max() > 1e-6:
                bad.append(("params", k, m, tuple(a), tuple(b)))
            n_tight += 1
    print(f"fit compare: {len(isl)} islands, {n_cmp} of >= 8 samples ({n_tight} of >= 12 compared by parameter); "
          f"worst cost rel diff {worst_cost:.2e}, worst parameter diff {worst_par:.2e}; "
          f"{len(other_minimum)} at another end point {other_minimum[:3]}; {n_unconverged} at the iteration limit; "
          f"{len(bad)} outside tolerance",
          flush=True)
    assert not bad, bad[:5]
    assert len(other_minimum) <= max(1, n_cmp // 100), other_minimum
    return n_tight


def test_gpu_fit_reference_kat(gpu, orc):
    truth = dict(x_mean=48.24, y_mean=52.66, amplitude=10.0, x_stddev=1.5, y_stddev=1.2, theta=0.5)
    img = np.asfortranarray(fftshift(gaussian2d(128, 64, **truth)))
    p = capi.sourcefind_pars(4.0, 3.0, rms_est=1.0, find_negative_sources=False, median_method=capi.MEDIAN_BINMEDIAN)
    sf = gpu.source_find(img, p)
    assert len(sf.islands) == 1
    f = gpu.fit_islands(img, sf.label_map, sf.islands)[0]
    want = [truth["amplitude"], truth["x_mean"], truth["y_mean"], truth["x_stddev"], truth["y_stddev"], truth["theta"]]
    assert np.allclose(f["leastsq_fit"], want, rtol=0, atol=1e-5), f
    o = orc.fit_islands(img, sf.label_map, list(sf.islands))[0]
    assert f["bounding_box"] == o["bounding_box"]
    assert np.allclose(f["leastsq_fit"], o["leastsq_fit"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32C])
def test_gpu_fit_c1_image(gpu, orc, precision):
    """Every island of an imaged c1-size snapshot (N=1024, 50k visibilities, 5 sources),
    fitted on the device image and label map, against the oracle on the same values."""
    snap = simdata.snapshot(1024, 1.0, 50_000, 5, seed=simdata.seed_for(1, 41))
    pars = capi.imager_pars(1024, 1.0)
    N = 1024
    nv = snap["weights"].size
    osz, dt = (8, np.float64) if precision == capi.PRECISION_FP64 else (4, np.float32)
    bufs = [gpu.malloc(16 * nv), gpu.malloc(8 * nv), gpu.malloc(24 * nv), gpu.malloc(osz * N * N),
            gpu.malloc(osz * N * N), gpu.malloc(4 * N * N)]
    try:
        gpu.htod(bufs[0], np.ascontiguousarray(snap["vis"]))
        gpu.htod(bufs[1], np.ascontiguousarray(snap["weights"]))
        gpu.htod(bufs[2], np.asfortranarray(snap["uvw"]))
        gpu.image_visibilities_device(pars, nv, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], precision=precision,
                                      synchronize=True)
        sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
        isl, _ = gpu.source_find_device(bufs[3], N, N, sfp, d_label_map=bufs[5], precision=precision)
        isl = list(isl)
        got = gpu.fit_islands_device(bufs[3], N, N, bufs[5], isl, precision=precision)
        img = np.empty((N, N), order="F", dtype=dt)
        lm = np.empty((N, N), order="F", dtype=np.int32)
        gpu.dtoh(img, bufs[3])
        gpu.dtoh(lm, bufs[5])
    finally:
        for b in bufs:
            gpu.free(b)
    img = img.astype(np.float64)
    want = orc.fit_islands(img, lm, isl)
    n_tight = _compare(orc, img, lm, isl, got, want)
    print(f"fit c1 precision {precision}: {len(isl)} islands, {n_tight} compared tightly", flush=True)
    assert len(isl) > 10 and n_tight > 0


def test_gpu_fit_noisy_sources(gpu, orc):
    """Elongated sources of both signs at assorted angles in unit noise: GPU and oracle
    agree, and each bright source's fit lies near its truth."""
    rng = np.random.default_rng(7)
    rows, cols = 256, 256
    img = rng.normal(size=(rows, cols))
    truths = []
    for k in range(12):
        t = np.array([rng.choice([-1, 1]) * rng.uniform(20, 60), rng.uniform(20, 230), rng.uniform(20, 230),
                      rng.uniform(1.5, 4.0), rng.uniform(0.8, 1.5), rng.uniform(-np.pi / 2, np.pi / 2)])
        img += gaussian2d(rows, cols, t[1], t[2], t[0], t[3], t[4], t[5])
        truths.append(t)
    img = np.asfortranarray(fftshift(img))
    p = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    sf = gpu.source_find(img, p)
    isl = list(sf.islands)
    got = gpu.fit_islands(img, sf.label_map, isl)
    want = orc.fit_islands(img, sf.label_map, isl)
    _compare(orc, img, sf.label_map, isl, got, want)
    for t in truths:
        best = min(got, key=lambda f: np.hypot(f["leastsq_fit"][1] - t[1], f["leastsq_fit"][2] - t[2]))
        assert np.hypot(best["leastsq_fit"][1] - t[1], best["leastsq_fit"][2] - t[2]) < 0.3
        assert abs(best["leastsq_fit"][0] - t[0]) < 0.15 * abs(t[0])


def test_gpu_fit_errors(gpu):
    img = np.zeros((16, 16), order="F")
    lm = np.zeros((16, 16), dtype=np.int32, order="F")
    lm[2, 2] = lm[2, 3] = lm[3, 2] = 1
    isl = [dict(label_idx=1, sign=1, extremum_y_idx=10, extremum_x_idx=10, num_samples=5, extremum_val=1.0,
                moments_fit=(1.0, 10.0, 10.0, 1.0, 1.0, 0.0))]
    from fastimaging_amd.stp import STPError
    with pytest.raises(STPError):
        gpu.fit_islands(img, lm, isl)  # 5 samples claimed, the label map holds 3
    assert gpu.fit_islands(img, lm, []) == []


def test_gpu_stp_python_gaussian_fitting():
    from fastimaging_amd import stp_python
    truth = dict(x_mean=48.24, y_mean=52.66, amplitude=10.0, x_stddev=1.5, y_stddev=1.2, theta=0.5)
    img = np.asfortranarray(fftshift(gaussian2d(128, 64, **truth)))
    out = stp_python.source_find_wrapper(img, 4.0, 3.0, rms_est=1.0, find_negative_sources=False,
                                         median_method=stp_python.MedianMethod.BINMEDIAN, gaussian_fitting=True)
    assert len(out) == 1
    lf = out[0][6]
    assert abs(lf.amplitude - 10.0) < 1e-5 and abs(lf.x_centre - 48.24) < 1e-5 and abs(lf.theta - 0.5) < 1e-5
