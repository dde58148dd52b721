"""Python mirror of the reference's stp:: interface over the HIP C-ABI.

    ctx = Context(0)
    image, beam = ctx.image_visibilities(vis, vis_weights, uvw_lambda, img_pars)   # stp::image_visibilities
    sf = ctx.source_find(image, pars)                                            # stp::SourceFindImage

Matrices are numpy arrays in Fortran (column-major) order, as the reference's
arma::Mat.  Errors raise STPError carrying the C status and message; there is
no fallback path: a missing library or device raises.
"""
import ctypes

import numpy as np

from . import capi

_P = ctypes.c_void_p


class STPError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{capi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


class IslandList:
    """Sequence view over the first n records of a capi.Island array; records
    are converted to dicts (capi.island_tuple) only when indexed."""

    def __init__(self, arr, n):
        self._arr = arr
        self._n = n

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self._n))]
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError(i)
        return capi.island_tuple(self._arr[i])

    def __iter__(self):
        return (capi.island_tuple(self._arr[i]) for i in range(self._n))

    def raw(self):
        """The underlying ctypes records (first len(self) are valid)."""
        return self._arr


class SourceFindResult:
    """Members of stp::SourceFindImage (sourcefind.h:146-164) that the C-ABI returns."""

    def __init__(self, label_map, islands, stats):
        self.label_map = label_map
        self.islands = islands
        self.bg_level = stats.bg_level
        self.rms_est = stats.rms_est
        self.stats = stats


class Context:
    def __init__(self, device=0):
        self.lib = capi.load()
        h = _P()
        st = self.lib.stp_ctx_create(ctypes.byref(h), device)
        if st != 0:
            raise STPError(st, f"stp_ctx_create(device={device}) failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.stp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise STPError(st, self.lib.stp_ctx_last_error(self.h).decode())

    # ----------------------------------------------------------------- imaging
    def image_visibilities(self, vis, vis_weights, uvw_lambda, img_pars, w_proj=None, a_proj=None,
                           precision=capi.PRECISION_FP64, return_diag=False, kernel=None):
        """stp::image_visibilities (imager.h:233-376); returns (image, beam).

        kernel: capi.KernelSpec (the kernel_creator functor) or None for the
        ImageVisibilities dispatch on img_pars (imager.cpp:26-60)."""
        uvw = np.asfortranarray(uvw_lambda, dtype=np.float64)
        n = uvw.shape[0]
        vis = np.ascontiguousarray(vis, dtype=np.complex128).reshape(n)
        w = np.ascontiguousarray(vis_weights, dtype=np.float64).reshape(n)
        N = img_pars.image_size
        image = np.empty((N, N), order="F")
        beam = np.empty((N, N), order="F") if img_pars.generate_beam else None
        diag = capi.ImagerDiag()
        self._check(self.lib.stp_image_visibilities(
            self.h, ctypes.byref(img_pars), ctypes.byref(w_proj) if w_proj is not None else None,
            ctypes.byref(a_proj) if a_proj is not None else None, ctypes.byref(kernel) if kernel is not None else None,
            n, _ptr(vis), _ptr(w), _ptr(uvw),
            precision, _ptr(image), _ptr(beam), ctypes.byref(diag)))
        if diag.image_rows == 0:
            image, beam = np.empty((0, 0)), (np.empty((0, 0)) if img_pars.generate_beam else None)
        return (image, beam, diag) if return_diag else (image, beam)

    def convolve_to_grid(self, kernel_function, kernel_param, support, image_size, uv, vis, weights,
                         kernel_exact=True, oversampling=1, shift_uv=True, halfplane=True, generate_beam=True,
                         widths=None):
        """stp::convolve_to_grid (gridder.h:240-997); returns (vis_grid, sampling_grid, total).

        kernel_function / kernel_param / widths describe the kernel_creator
        functor (widths = (gaussian, sinc) width normalisations, None for the
        functor's defaults)."""
        wg, ws = widths if widths is not None else (None, None)
        spec = capi.kernel_spec(kernel_function, float(kernel_param), wg, ws)
        uv = np.asfortranarray(uv, dtype=np.float64)
        n = uv.shape[0]
        vis = np.ascontiguousarray(vis, dtype=np.complex128).reshape(n)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(n)
        rows = image_size // 2 + 1 if halfplane else image_size
        vg = np.empty((rows, image_size), dtype=np.complex128, order="F")
        sg = np.empty((rows, image_size), dtype=np.complex128, order="F") if generate_beam else None
        total = ctypes.c_double()
        self._check(self.lib.stp_convolve_to_grid(
            self.h, ctypes.byref(spec), support, image_size, n, _ptr(uv), _ptr(vis), _ptr(w),
            int(kernel_exact), oversampling, int(shift_uv), int(halfplane), int(generate_beam), _ptr(vg), _ptr(sg),
            ctypes.byref(total)))
        return vg, sg, total.value

    def fft_c2r(self, grid):
        """stp::fft_fftw_c2r (fft.cpp:45-135)."""
        g = np.asfortranarray(grid, dtype=np.complex128)
        in_rows, cols = g.shape
        n = 2 * (in_rows - 1)
        out = np.empty((n, cols), order="F")
        self._check(self.lib.stp_fft_c2r(self.h, _ptr(g), in_rows, cols, _ptr(out)))
        return out

    def fft_c2c(self, a, forward=True):
        """stp::fft_fftw_c2c (fft.cpp:227-323), square matrices."""
        a = np.asfortranarray(a, dtype=np.complex128)
        out = np.empty_like(a, order="F")
        self._check(self.lib.stp_fft_c2c(self.h, _ptr(a), a.shape[0], int(forward), _ptr(out)))
        return out

    # ------------------------------------------------------------ source find
    def generate_visibilities_from_local_skymodel(self, skymodel, uvw_lambda):
        """stp::generate_visibilities_from_local_skymodel (visibility.cpp:31-41): skymodel
        rows (l, m, flux), uvw (n, 3) in wavelengths; returns complex128 (n,)."""
        sm = np.asfortranarray(skymodel, dtype=np.float64).reshape(-1, 3)
        uvw = np.asfortranarray(uvw_lambda, dtype=np.float64)
        out = np.empty(uvw.shape[0], dtype=np.complex128)
        self._check(self.lib.stp_generate_visibilities_from_local_skymodel(self.h, sm.shape[0], _ptr(sm), uvw.shape[0],
                                                                           _ptr(uvw), _ptr(out)))
        return out

    def model_visibilities_device(self, skymodel, n_vis, d_uvw, d_vis, subtract=True, synchronize=False):
        """Model (subtract=False) or residual (d_vis -= model) visibilities on device buffers."""
        sm = np.asfortranarray(skymodel, dtype=np.float64).reshape(-1, 3)
        self._check(self.lib.stp_model_visibilities_device(self.h, sm.shape[0], _ptr(sm), n_vis, d_uvw, d_vis,
                                                           int(subtract), int(synchronize)))

    def source_find(self, image, pars, capacity=1 << 16, label_map=True):
        """stp::SourceFindImage (sourcefind.cpp:156-292)."""
        img = np.asfortranarray(image, dtype=np.float64)
        rows, cols = img.shape
        lm = np.empty((rows, cols), dtype=np.int32, order="F") if label_map else None
        isl = (capi.Island * capacity)()
        n = ctypes.c_size_t()
        st = capi.SourceFindStats()
        self._check(self.lib.stp_source_find(self.h, _ptr(img), rows, cols, ctypes.byref(pars), _ptr(lm), isl,
                                             capacity, ctypes.byref(n), ctypes.byref(st)))
        return SourceFindResult(lm, [capi.island_tuple(isl[i]) for i in range(n.value)], st)

    # ------------------------------------------------------- device buffers
    def fit_islands(self, image, label_map, islands):
        """IslandParams::leastsq_fit_gaussian_2d (sourcefind.cpp:719-866) for each island
        of a source find (dicts or capi.Island records) on its image and label map;
        returns a list of dicts (leastsq_fit, bounding_box, iterations, status)."""
        img = np.asfortranarray(image, dtype=np.float64)
        lm = np.asfortranarray(label_map, dtype=np.int32)
        rows, cols = img.shape
        islands = list(islands.raw()[:len(islands)] if isinstance(islands, IslandList) else islands)
        n = len(islands)
        recs = capi.islands_to_records(islands)
        out = (capi.FitResult * max(1, n))()
        self._check(self.lib.stp_fit_islands(self.h, _ptr(img), rows, cols, _ptr(lm), ctypes.cast(recs, _P), n,
                                             ctypes.cast(out, _P)))
        return [capi.fit_tuple(out[i]) for i in range(n)]

    def fit_islands_device(self, d_image, rows, cols, d_label_map, islands, precision=0):
        """The same on a device image (element type by precision) and device label map."""
        islands = list(islands.raw()[:len(islands)] if isinstance(islands, IslandList) else islands)
        n = len(islands)
        recs = capi.islands_to_records(islands)
        out = (capi.FitResult * max(1, n))()
        self._check(self.lib.stp_fit_islands_device(self.h, d_image, precision, rows, cols, d_label_map,
                                                    ctypes.cast(recs, _P), n, ctypes.cast(out, _P)))
        return [capi.fit_tuple(out[i]) for i in range(n)]

    def malloc(self, nbytes):
        p = _P()
        self._check(self.lib.stp_device_malloc(self.h, ctypes.byref(p), nbytes))
        return p

    def free(self, p):
        self._check(self.lib.stp_device_free(self.h, p))

    def island_buffer(self, capacity):
        """A (capi.Island * capacity) array in page-locked host memory: island
        lists written to it arrive by DMA (free with free_host)."""
        p = _P()
        self._check(self.lib.stp_host_malloc(self.h, ctypes.byref(p), ctypes.sizeof(capi.Island) * capacity))
        arr = (capi.Island * capacity).from_address(p.value)
        arr._stp_host_ptr = p
        return arr

    def free_host(self, arr):
        self._check(self.lib.stp_host_free(self.h, arr._stp_host_ptr))

    def htod(self, dptr, arr):
        arr = np.ascontiguousarray(arr) if not arr.flags.f_contiguous else arr
        self._check(self.lib.stp_memcpy_htod(self.h, dptr, _ptr(arr), arr.nbytes))

    def dtoh(self, arr, dptr):
        self._check(self.lib.stp_memcpy_dtoh(self.h, _ptr(arr), dptr, arr.nbytes))

    def image_visibilities_device(self, img_pars, n, d_vis, d_w, d_uvw, d_image, d_beam, precision=0,
                                  synchronize=False, w_proj=None, kernel=None, noalloc=False):
        """Imaging on device buffers.  noalloc=True calls the entry point that
        takes every intermediate buffer from the library's fixed device scratch
        region (capi.NOALLOC_SCRATCH_BYTES) and allocates nothing."""
        diag = capi.ImagerDiag()
        fn = self.lib.stp_image_visibilities_device_noalloc if noalloc else self.lib.stp_image_visibilities_device
        self._check(fn(
            self.h, ctypes.byref(img_pars), ctypes.byref(w_proj) if w_proj is not None else None, None,
            ctypes.byref(kernel) if kernel is not None else None, n,
            d_vis, d_w, d_uvw, precision, d_image, d_beam, int(synchronize),
            ctypes.byref(diag) if synchronize else None))
        return diag

    def source_find_device(self, d_image, rows, cols, pars, d_label_map=None, precision=0, capacity=1 << 16,
                           out=None):
        """Source finding on a device-resident image.  Returns (IslandList, stats).
        out: a reusable (capi.Island * capacity) array; the returned list then
        views it and is overwritten by the next call that reuses it."""
        if out is not None:
            isl, capacity = out, len(out)
        else:
            isl = (capi.Island * capacity)()
        n = ctypes.c_size_t()
        st = capi.SourceFindStats()
        self._check(self.lib.stp_source_find_device(self.h, d_image, precision, rows, cols, ctypes.byref(pars),
                                                    d_label_map, isl, capacity, ctypes.byref(n), ctypes.byref(st)))
        return IslandList(isl, n.value), st

    # ------------------------------------------------------- batches / tuning
    def set_autotune(self, mode=1, tune_file=None):
        self._check(self.lib.stp_ctx_set_autotune(self.h, mode, tune_file.encode() if tune_file else None))

    def get_tuning(self, img_pars, n_vis, precision=capi.PRECISION_FP64, w_proj=None):
        tc, l1, us = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        self._check(self.lib.stp_ctx_get_tuning(self.h, ctypes.byref(img_pars),
                                                ctypes.byref(w_proj) if w_proj is not None else None, n_vis, precision,
                                                ctypes.byref(tc), ctypes.byref(l1), ctypes.byref(us)))
        return tc.value, l1.value, us.value

    def image_batch(self, items, precision=capi.PRECISION_FP64, max_concurrent=0, flags=0, kernel=None):
        """stp_image_batch over capi.BatchItem records; returns the per-snapshot sample_grid_total."""
        arr = (capi.BatchItem * len(items))(*items)
        totals = np.zeros(len(items))
        self._check(self.lib.stp_image_batch(self.h, len(items), arr,
                                             ctypes.byref(kernel) if kernel is not None else None, precision,
                                             max_concurrent, flags, totals.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return totals

    def synchronize(self):
        self._check(self.lib.stp_ctx_synchronize(self.h))

    # ----------------------------------------------------------------- timing
    def set_timing(self, enable=True):
        self._check(self.lib.stp_ctx_set_timing(self.h, int(enable)))

    def set_timing_filter(self, names=None):
        """Times only the named kernels (iterable of names); None times every launch."""
        arg = ",".join(names).encode() if names else None
        self._check(self.lib.stp_ctx_set_timing_filter(self.h, arg))

    def reset_timings(self):
        self._check(self.lib.stp_ctx_reset_timings(self.h))

    def timings(self):
        """{kernel name: (total ms, launches)} from HIP events on the context stream."""
        cap = 64
        names = ctypes.create_string_buffer(64 * cap)
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_uint64 * cap)()
        n = ctypes.c_size_t()
        self._check(self.lib.stp_ctx_get_timings(self.h, names, ms, cnt, cap, ctypes.byref(n)))
        out = {}
        for i in range(min(n.value, cap)):
            nm = names.raw[64 * i: 64 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (ms[i], cnt[i])
        return out
