"""ORACLE - TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/libstp_oracle.so, the FP64 CPU restatement of the
reference Slow Transients Pipeline path (see oracle/oracle.h).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os

import numpy as np

from fastimaging_amd import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstp_oracle.so")

_lib = None
_P = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_last_error.restype = ctypes.c_char_p
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _spec(kind, param, widths):
    """KernelSpec of a functor; widths = (width_normalization_gaussian,
    width_normalization_sinc) or None for the reference's defaults."""
    wg, ws = widths if widths is not None else (None, None)
    return ctypes.byref(capi.kernel_spec(kind, param, wg, ws))


def threads():
    return lib().oracle_default_threads()


def kernel_eval(kind, param, x, widths=None):
    x = _f64(x)
    out = np.empty_like(x)
    _check(lib().oracle_kernel_eval(_spec(kind, param, widths), _ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return out


def make_1d_kernel(kind, param, support, offset, oversampling=1, pad=False, normalize=True, widths=None):
    n = 2 * (support + (1 if pad else 0)) * oversampling + 1
    out = np.empty(n)
    _check(lib().oracle_make_1d_kernel(_spec(kind, param, widths), ctypes.c_int(support),
                                       ctypes.c_double(offset), ctypes.c_int(oversampling), ctypes.c_int(int(pad)),
                                       ctypes.c_int(int(normalize)), _ptr(out)))
    return out


def make_kernel_array(kind, param, support, offset, oversampling=1, pad=False, normalize=True, widths=None):
    n = 2 * (support + (1 if pad else 0)) * oversampling + 1
    out = np.empty((n, n), order="F")
    _check(lib().oracle_make_kernel_array(_spec(kind, param, widths), ctypes.c_int(support),
                                          ctypes.c_double(offset[0]), ctypes.c_double(offset[1]),
                                          ctypes.c_int(oversampling), ctypes.c_int(int(pad)),
                                          ctypes.c_int(int(normalize)), _ptr(out)))
    return out


def img_dom_kernel(kind, param, size, normalize=False, analytic=False, widths=None):
    out = np.empty(size)
    _check(lib().oracle_img_dom_kernel(_spec(kind, param, widths), ctypes.c_size_t(size),
                                       ctypes.c_int(int(normalize)), ctypes.c_int(int(analytic)), _ptr(out)))
    return out


def oversampled_indices(frac, oversampling):
    frac = _f64(frac)
    out = np.empty(frac.shape, dtype=np.int32)
    _check(lib().oracle_oversampled_indices(_ptr(frac), ctypes.c_size_t(frac.size), ctypes.c_uint(oversampling), _ptr(out)))
    return out


def convolve_to_grid(kind, param, support, image_size, uv, vis, weights, kernel_exact=True, oversampling=1,
                     shift_uv=True, halfplane=True, generate_beam=True, wproj=None, w_lambda=None,
                     cell_size=0.0, analytic_gcf=True, nthreads=None, widths=None):
    """convolve_to_grid (src/stp/gridder/gridder.h:240-997). uv: (n, 2) pixel units."""
    uv = np.asfortranarray(uv, dtype=np.float64)
    n = uv.shape[0]
    vis = np.ascontiguousarray(vis, dtype=np.complex128).reshape(n)
    weights = _f64(weights).reshape(n)
    rows = image_size // 2 + 1 if halfplane else image_size
    vg = np.zeros((rows, image_size), dtype=np.complex128, order="F")
    sg = np.zeros((rows, image_size), dtype=np.complex128, order="F") if generate_beam else None
    total = ctypes.c_double()
    ngood = ctypes.c_uint64()
    wl = _f64(w_lambda) if w_lambda is not None else None
    wp = wproj if wproj is not None else None
    _check(lib().oracle_convolve_to_grid(
        _spec(kind, param, widths), ctypes.c_uint(support), ctypes.c_int(image_size),
        ctypes.c_size_t(n), _ptr(uv), _ptr(vis), _ptr(weights), ctypes.c_int(int(kernel_exact)),
        ctypes.c_uint(oversampling), ctypes.c_int(int(shift_uv)), ctypes.c_int(int(halfplane)),
        ctypes.c_int(int(generate_beam)), ctypes.byref(wp) if wp is not None else None, _ptr(wl),
        ctypes.c_double(cell_size), ctypes.c_int(int(analytic_gcf)), ctypes.c_int(nthreads or threads()),
        _ptr(vg), _ptr(sg), ctypes.byref(total), ctypes.byref(ngood)))
    return vg, sg, total.value, ngood.value


def fft_c2r(grid, nthreads=None):
    g = np.asfortranarray(grid, dtype=np.complex128)
    in_rows, cols = g.shape
    n = 2 * (in_rows - 1)
    out = np.empty((n, cols), order="F")
    _check(lib().oracle_fft_c2r(_ptr(g), ctypes.c_int(in_rows), ctypes.c_int(cols), _ptr(out),
                                ctypes.c_int(nthreads or threads())))
    return out


def fft_c2c(a, forward=True, nthreads=None):
    a = np.asfortranarray(a, dtype=np.complex128)
    out = np.empty_like(a, order="F")
    _check(lib().oracle_fft_c2c(_ptr(a), ctypes.c_int(a.shape[0]), ctypes.c_int(a.shape[1]),
                                ctypes.c_int(int(forward)), _ptr(out), ctypes.c_int(nthreads or threads())))
    return out


def wproj_kernel_cache(w_value, aa, workarea, cell_rad, oversampling, scaling, support, nthreads=None,
                       trunc_perc=0.0, return_support=False):
    """generate_kernel_cache (aw_projection.cpp:366-452) of one plane; with trunc_perc > 0
    the support is truncated first (aw_projection.cpp:197-209) and the cache is that size."""
    cs = (oversampling // 2) * 2 + 1
    ks = 2 * support + 1
    out = np.empty((cs, cs, ks, ks), dtype=np.complex128)  # [y][x] then (col-major ks x ks), room for the max
    aa = _f64(aa)
    ts = ctypes.c_uint()
    _check(lib().oracle_wproj_kernel_cache(ctypes.c_double(w_value), _ptr(aa), ctypes.c_uint(workarea),
                                           ctypes.c_double(cell_rad), ctypes.c_uint(oversampling),
                                           ctypes.c_double(scaling), ctypes.c_uint(support),
                                           ctypes.c_int(nthreads or threads()), ctypes.c_double(trunc_perc),
                                           ctypes.byref(ts), _ptr(out)))
    kt = 2 * ts.value + 1
    flat = out.reshape(-1)[: cs * cs * kt * kt].reshape(cs, cs, kt, kt)
    # out[y, x] holds a kt x kt column-major matrix -> transpose the last two axes
    res = np.ascontiguousarray(np.swapaxes(flat, 2, 3))
    return (res, ts.value) if return_support else res


def hankel_kernel_cache(w_value, aa, workarea, cell_rad, oversampling, scaling, support, proj_slice=False, cubic=False,
                        trunc_perc=0.0):
    """Hankel W-kernel cache of one plane (aw_projection.cpp:211-287,380-411): returns
    (cache[y, x, row, col], truncated support, transformed radial profile)."""
    cs = (oversampling // 2) * 2 + 1
    ks = 2 * support + 1
    out = np.empty((cs, cs, ks, ks), dtype=np.complex128)
    aa = _f64(aa)
    ts = ctypes.c_uint()
    cap = workarea * oversampling
    rad = np.empty(cap, dtype=np.complex128)
    rn = ctypes.c_size_t()
    _check(lib().oracle_hankel_kernel_cache(ctypes.c_double(w_value), _ptr(aa), ctypes.c_uint(workarea),
                                            ctypes.c_double(cell_rad), ctypes.c_uint(oversampling),
                                            ctypes.c_double(scaling), ctypes.c_uint(support), ctypes.c_int(int(proj_slice)),
                                            ctypes.c_int(int(cubic)), ctypes.c_double(trunc_perc), ctypes.byref(ts),
                                            _ptr(out), _ptr(rad), ctypes.c_size_t(cap), ctypes.byref(rn)))
    kt = 2 * ts.value + 1
    flat = out.reshape(-1)[: cs * cs * kt * kt].reshape(cs, cs, kt, kt)
    return np.ascontiguousarray(np.swapaxes(flat, 2, 3)), ts.value, rad[: rn.value].copy()


def besselj1(x):
    """besselj1 (aw_projection.cpp:490-508)."""
    x = _f64(np.atleast_1d(x))
    out = np.empty_like(x)
    _check(lib().oracle_besselj1(_ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return out


def dht_matrix(n):
    """dht (aw_projection.cpp:530-596): D[j, i] (numpy row j, column i)."""
    out = np.empty(n * n)
    _check(lib().oracle_dht_matrix(ctypes.c_size_t(n), _ptr(out)))
    return out.reshape(n, n).T.copy()  # column-major D[j + i n]


def spline(y, x, cubic=True):
    """tk::spline (natural cubic or linear) through (k, y[k]), k = 0..n-1, evaluated at x."""
    y = _f64(y)
    x = _f64(np.atleast_1d(x))
    out = np.empty_like(x)
    _check(lib().oracle_spline(_ptr(y), ctypes.c_size_t(y.size), ctypes.c_int(int(cubic)), _ptr(x), ctypes.c_size_t(x.size),
                               _ptr(out)))
    return out


def average_w_planes(abs_w_sorted, good, planes, median=False):
    """average_w_planes (gridder.cpp:68-146) on |w| already in ascending order."""
    w = _f64(abs_w_sorted)
    g = np.ascontiguousarray(good, dtype=np.uint32)
    avg = np.empty(planes)
    first = np.empty(planes, dtype=np.uint64)
    _check(lib().oracle_average_w_planes(_ptr(w), _ptr(g), ctypes.c_size_t(w.size), ctypes.c_uint(planes),
                                         ctypes.c_int(int(median)), _ptr(avg), _ptr(first)))
    return avg, first


def image_visibilities(pars, vis, weights, uvw, wproj=None, nthreads=None, kernel=None):
    """image_visibilities (src/stp/imager/imager.h:233-376). Returns (image, beam, total, ngood)."""
    uvw = np.asfortranarray(uvw, dtype=np.float64)
    n = uvw.shape[0]
    vis = np.ascontiguousarray(vis, dtype=np.complex128).reshape(n)
    weights = _f64(weights).reshape(n)
    N = pars.image_size
    image = np.empty((N, N), order="F")
    beam = np.empty((N, N), order="F") if pars.generate_beam else None
    total = ctypes.c_double()
    ngood = ctypes.c_uint64()
    empty = ctypes.c_int()
    _check(lib().oracle_image_visibilities(ctypes.byref(pars), ctypes.byref(wproj) if wproj is not None else None,
                                           ctypes.byref(kernel) if kernel is not None else None,
                                           ctypes.c_size_t(n), _ptr(vis), _ptr(weights), _ptr(uvw),
                                           ctypes.c_int(nthreads or threads()), _ptr(image), _ptr(beam),
                                           ctypes.byref(total), ctypes.byref(ngood), ctypes.byref(empty)))
    if empty.value:
        return None, None, total.value, ngood.value
    return image, beam, total.value, ngood.value


def canonical_pair_sum(x):
    x = _f64(np.ravel(x, order="F"))
    out = np.empty(2)
    _check(lib().oracle_canonical_pair_sum(_ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return out[0], out[1]


def mean_stddev(x):
    x = _f64(np.ravel(x, order="F"))
    out = np.empty(2)
    _check(lib().oracle_mean_stddev(_ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return out[0], out[1]


def median_exact(x):
    x = _f64(np.ravel(x, order="F"))
    out = ctypes.c_double()
    _check(lib().oracle_median_exact(_ptr(x), ctypes.c_size_t(x.size), ctypes.byref(out)))
    return out.value


def binmedian(x):
    x = _f64(np.ravel(x, order="F"))
    out = np.empty(3)
    _check(lib().oracle_binmedian(_ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return tuple(out)


def binapprox(x):
    x = _f64(np.ravel(x, order="F"))
    out = np.empty(3)
    _check(lib().oracle_binapprox(_ptr(x), ctypes.c_size_t(x.size), _ptr(out)))
    return tuple(out)


def estimate_rms(x, num_sigma=3.0, iters=5, stats=None):
    """estimate_rms (sourcefind.cpp:22-154). stats: (mean, sigma, median, valid_bits) or None."""
    x = _f64(np.ravel(x, order="F"))
    if stats is None:
        s3, valid = np.zeros(3), 0
    else:
        s3, valid = np.array(stats[:3], dtype=np.float64), int(stats[3])
    out = ctypes.c_double()
    _check(lib().oracle_estimate_rms(_ptr(x), ctypes.c_size_t(x.size), ctypes.c_double(num_sigma),
                                     ctypes.c_uint(iters), _ptr(s3), ctypes.c_int(valid), ctypes.byref(out)))
    return out.value


def ccl(img, thr_pos, thr_neg, find_negative=True, conn4=False, sauf_slab_cols=None):
    img = np.asfortranarray(img, dtype=np.float64)
    rows, cols = img.shape
    labels = np.zeros((rows, cols), dtype=np.int32, order="F")
    npos, nneg = ctypes.c_uint32(), ctypes.c_uint32()
    if sauf_slab_cols is None:
        _check(lib().oracle_ccl(_ptr(img), ctypes.c_int(rows), ctypes.c_int(cols), ctypes.c_double(thr_pos),
                                ctypes.c_double(thr_neg), ctypes.c_int(int(find_negative)), ctypes.c_int(int(conn4)),
                                _ptr(labels), ctypes.byref(npos), ctypes.byref(nneg)))
    else:
        _check(lib().oracle_ccl_sauf(_ptr(img), ctypes.c_int(rows), ctypes.c_int(cols), ctypes.c_double(thr_pos),
                                     ctypes.c_double(thr_neg), ctypes.c_int(int(find_negative)),
                                     ctypes.c_int(int(conn4)), ctypes.c_int(sauf_slab_cols), _ptr(labels),
                                     ctypes.byref(npos), ctypes.byref(nneg)))
    return labels, npos.value, nneg.value


def source_find(img, pars, capacity=100000):
    """SourceFindImage (sourcefind.cpp:156-292). Returns (label_map, islands(list of dict), stats)."""
    img = np.asfortranarray(img, dtype=np.float64)
    rows, cols = img.shape
    lm = np.zeros((rows, cols), dtype=np.int32, order="F")
    isl = (capi.Island * capacity)()
    n = ctypes.c_size_t()
    st = capi.SourceFindStats()
    _check(lib().oracle_source_find(_ptr(img), ctypes.c_int(rows), ctypes.c_int(cols), ctypes.byref(pars),
                                    _ptr(lm), isl, ctypes.c_size_t(capacity), ctypes.byref(n), ctypes.byref(st)))
    if n.value > capacity:
        raise RuntimeError("island capacity exceeded")
    return lm, [capi.island_tuple(isl[i]) for i in range(n.value)], st


def populate_kernel_cache(kind, param, support, oversampling, widths=None):
    """populate_kernel_cache (gridder.h:73-105): returns [a, b] -> (2s+1, 2s+1) real matrices."""
    cs = (oversampling // 2) * 2 + 1
    ks = 2 * support + 1
    out = np.empty((cs, cs, ks, ks), dtype=np.complex128)
    _check(lib().oracle_populate_kernel_cache(_spec(kind, param, widths), ctypes.c_uint(support),
                                              ctypes.c_uint(oversampling), _ptr(out)))
    return np.ascontiguousarray(np.swapaxes(out, 2, 3))


def model_visibilities(skymodel, uvw, nthreads=None):
    """generate_visibilities_from_local_skymodel (visibility.cpp:31-41): skymodel (n_src, 3)
    rows (l, m, flux), uvw (n_vis, 3) in wavelengths; returns complex128 (n_vis,)."""
    sm = np.asfortranarray(skymodel, dtype=np.float64).reshape(-1, 3)
    uvw = np.asfortranarray(uvw, dtype=np.float64)
    out = np.zeros(uvw.shape[0], dtype=np.complex128)
    _check(lib().oracle_model_visibilities(_ptr(sm), ctypes.c_size_t(sm.shape[0]), _ptr(uvw),
                                           ctypes.c_size_t(uvw.shape[0]), _ptr(out),
                                           nthreads if nthreads is not None else threads()))
    return out


def gaussian_jacobian(params, x, y, v):
    """Residual model - data and its Jacobian row (fitting.cpp:59-120); params =
    (amplitude, x_centre, y_centre, x_stddev, y_stddev, theta)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    r = ctypes.c_double()
    J = np.zeros(6)
    _check(lib().oracle_gaussian_jacobian(_ptr(p), ctypes.c_double(x), ctypes.c_double(y), ctypes.c_double(v),
                                          ctypes.byref(r), _ptr(J)))
    return r.value, J


def lm_fit(xs, ys, vs, start):
    """The island fit's minimiser on samples (xs, ys, vs) from `start`; returns
    (constrained params (6,), half cost, iterations, status)."""
    xs, ys, vs = (np.ascontiguousarray(a, dtype=np.float64) for a in (xs, ys, vs))
    st = np.ascontiguousarray(start, dtype=np.float64)
    out = np.zeros(6)
    hc = ctypes.c_double()
    it = ctypes.c_int()
    status = ctypes.c_int()
    _check(lib().oracle_lm_fit(_ptr(xs), _ptr(ys), _ptr(vs), ctypes.c_size_t(xs.size), _ptr(st), _ptr(out),
                               ctypes.byref(hc), ctypes.byref(it), ctypes.byref(status)))
    return out, hc.value, it.value, status.value


def fit_islands(img, label_map, islands):
    """IslandParams::leastsq_fit_gaussian_2d (sourcefind.cpp:719-866) for each island
    (a list of capi.island_tuple dicts or capi.Island records); returns a list of
    dicts (leastsq_fit, bounding_box, iterations, status)."""
    img = np.asfortranarray(img, dtype=np.float64)
    lm = np.asfortranarray(label_map, dtype=np.int32)
    rows, cols = img.shape
    islands = list(islands)
    n = len(islands)
    recs = capi.islands_to_records(islands)
    out = (capi.FitResult * max(1, n))()
    _check(lib().oracle_fit_islands(_ptr(img), ctypes.c_int(rows), ctypes.c_int(cols), _ptr(lm), recs,
                                    ctypes.c_size_t(n), out))
    return [capi.fit_tuple(out[i]) for i in range(n)]
