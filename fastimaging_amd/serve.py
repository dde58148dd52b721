"""Python side of the serving loop (C-ABI stp_server_*): submit snapshot imaging +
source-finding jobs as they arrive, wait for results.

    srv = Server(num_slots=4, hbm_budget_bytes=64 << 30)
    jid = srv.submit(vis, weights, uvw, img_pars, sourcefind=sf_pars)
    out = srv.wait(jid)          # dict: image, beam, islands, stats, diag, timings

Each job's numpy buffers are held by the Server until its result is collected.
"""
import ctypes

import numpy as np

from . import capi
from .stp import STPError, _ptr


class Server:
    def __init__(self, num_slots=4, hbm_budget_bytes=0, device=0):
        self.lib = capi.load()
        cfg = capi.ServerCfg(device, num_slots, hbm_budget_bytes)
        h = ctypes.c_void_p()
        st = self.lib.stp_server_create(ctypes.byref(h), ctypes.byref(cfg))
        if st != 0:
            raise STPError(st, "stp_server_create failed")
        self.h = h
        self._held = {}

    def close(self):
        if self.h:
            self.lib.stp_server_destroy(self.h)
            self.h = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, vis, vis_weights, uvw_lambda, img_pars, w_proj=None, kernel=None,
               precision=capi.PRECISION_FP64, sourcefind=None, label_map=False, island_cap=None,
               want_image=True, want_beam=True):
        uvw = np.asfortranarray(uvw_lambda, dtype=np.float64)
        n = uvw.shape[0]
        vis = np.ascontiguousarray(vis, dtype=np.complex128).reshape(n)
        w = np.ascontiguousarray(vis_weights, dtype=np.float64).reshape(n)
        N = img_pars.image_size
        if island_cap is None:
            # one island per 256 pixels, at least 65536 (an 8192^2 image of the served
            # mix has held 92 000 islands)
            island_cap = max(1 << 16, N * N // 256)
        image = np.zeros((N, N), order="F") if want_image else None
        beam = np.zeros((N, N), order="F") if (want_beam and img_pars.generate_beam) else None
        lab = np.zeros((N, N), dtype=np.int32, order="F") if (label_map and sourcefind is not None) else None
        isl = (capi.Island * island_cap)() if sourcefind is not None else None
        job = capi.Job()
        job.img_pars = img_pars
        if w_proj is not None:
            job.w_proj = w_proj
        if kernel is not None:
            job.use_kernel_spec = 1
            job.kernel = kernel
        job.precision = precision
        job.n_vis = n
        job.vis, job.vis_weights, job.uvw_lambda = _ptr(vis), _ptr(w), _ptr(uvw)
        job.image = _ptr(image)
        job.beam = _ptr(beam)
        if sourcefind is not None:
            job.source_find = 1
            job.sf_pars = sourcefind
            job.label_map = _ptr(lab)
            job.islands = ctypes.cast(isl, ctypes.c_void_p)
            job.island_cap = island_cap
        jid = ctypes.c_uint64()
        st = self.lib.stp_server_submit(self.h, ctypes.byref(job), ctypes.byref(jid))
        if st != 0:
            raise STPError(st, self.lib.stp_server_last_error(self.h).decode())
        self._held[jid.value] = (vis, w, uvw, image, beam, lab, isl)
        return jid.value

    def poll(self, jid):
        d = ctypes.c_int32()
        st = self.lib.stp_server_poll(self.h, jid, ctypes.byref(d))
        if st != 0:
            raise STPError(st, "unknown job")
        return bool(d.value)

    def wait(self, jid):
        r = capi.JobResult()
        st = self.lib.stp_server_wait(self.h, jid, ctypes.byref(r))
        if st != 0:
            raise STPError(st, self.lib.stp_server_last_error(self.h).decode())
        vis, w, uvw, image, beam, lab, isl = self._held.pop(jid)
        if r.status != 0:
            raise STPError(r.status, r.error.decode())
        nis = min(r.n_islands, len(isl)) if isl is not None else 0
        empty = r.diag.image_rows == 0
        return {
            "image": None if image is None else (np.empty((0, 0)) if empty else image),
            "beam": None if beam is None else (np.empty((0, 0)) if empty else beam),
            "label_map": lab,
            "islands": [capi.island_tuple(isl[i]) for i in range(nis)] if isl is not None else None,
            "stats": r.stats,
            "diag": r.diag,
            "slot": r.slot,
            "t_submit": r.t_submit, "t_start": r.t_start, "t_end": r.t_end,
            "job_bytes": r.job_bytes, "slot_bytes_before": r.slot_bytes_before,
            "slot_bytes_after": r.slot_bytes_after,
        }

    def job_bytes(self, img_pars, n_vis, precision=capi.PRECISION_FP64, w_proj=None):
        job = capi.Job()
        job.img_pars = img_pars
        job.n_vis = n_vis
        job.precision = precision
        if w_proj is not None:
            job.w_proj = w_proj
        return int(self.lib.stp_job_device_bytes(ctypes.byref(job)))

    def load(self):
        q, r, b = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        self.lib.stp_server_load(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b))
        return q.value, r.value, b.value
