"""GPU tests of batches (stp_image_batch) and the autotuner: a batch of snapshots
with uneven sizes gives the images each snapshot gives alone (FP64: within 1e-12 of
the peak, the order of the LDS atomics being the only difference), every replay of
the captured hipGraph matches the eager run, and autotuned launch configurations
keep the oracle parity."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu


class _Batch:
    """Snapshots with their inputs and outputs in device memory."""

    def __init__(self, gpu, specs, precision):
        self.gpu, self.prec = gpu, precision
        self.snaps, self.items, self.bufs = [], [], []
        osz = 8 if precision == capi.PRECISION_FP64 else 4
        for spec in specs:
            N, cell, nv, ns, seed, wp, wmax = spec[:7]
            kw = spec[7] if len(spec) > 7 else {}
            snap = simdata.snapshot(N, cell, nv, ns, seed=seed, field_frac=0.5 if wp is not None else 0.25, w_max=wmax)
            b = [gpu.malloc(16 * nv), gpu.malloc(8 * nv), gpu.malloc(24 * nv), gpu.malloc(osz * N * N),
                 gpu.malloc(osz * N * N)]
            gpu.htod(b[0], np.ascontiguousarray(snap["vis"]))
            gpu.htod(b[1], np.ascontiguousarray(snap["weights"]))
            gpu.htod(b[2], np.asfortranarray(snap["uvw"]))
            it = capi.BatchItem()
            it.img_pars = capi.imager_pars(N, cell, **kw)
            if wp is not None:
                it.w_proj = wp
            it.n_vis = nv
            it.d_vis, it.d_vis_weights, it.d_uvw_lambda, it.d_image, it.d_beam = [x.value for x in b]
            self.snaps.append(snap)
            self.items.append(it)
            self.bufs.append(b)

    def images(self):
        out = []
        dt = np.float64 if self.prec == capi.PRECISION_FP64 else np.float32
        for it in self.items:
            N = it.img_pars.image_size
            img = np.empty((N, N), order="F", dtype=dt)
            beam = np.empty((N, N), order="F", dtype=dt)
            self.gpu.dtoh(img, it.d_image)
            self.gpu.dtoh(beam, it.d_beam)
            out.append((img.astype(np.float64), beam.astype(np.float64)))
        return out

    def clear(self):
        for it in self.items:
            N = it.img_pars.image_size
            z = np.zeros(N * N * (8 if self.prec == capi.PRECISION_FP64 else 4), dtype=np.uint8)
            self.gpu.htod(it.d_image, z)
            self.gpu.htod(it.d_beam, z)

    def free(self):
        for b in self.bufs:
            for x in b:
                self.gpu.free(x)


SPECS = [
    (512, 2.0, 20_000, 5, 701, None, 0.0),
    (1024, 1.0, 120_000, 12, 702, None, 0.0),
    (256, 4.0, 5_000, 3, 703, None, 0.0),
    (1024, 1.0, 60_000, 30, 704, None, 0.0),
    (256, 20.0, 4_000, 6, 705, "w", 300.0),
    (2048, 1.0, 300_000, 40, 706, None, 0.0),
]


def _specs():
    return [(N, c, nv, ns, sd, capi.wproj_pars(4, 8) if wp else None, wm) for N, c, nv, ns, sd, wp, wm in SPECS]


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32C])
def test_batch_matches_single_and_graph_replays(gpu, orc, precision):
    b = _Batch(gpu, _specs(), precision)
    try:
        tol = 1e-12 if precision == capi.PRECISION_FP64 else 1e-6
        # each snapshot alone through the single-snapshot entry point
        single = []
        for it in b.items:
            gpu.image_visibilities_device(it.img_pars, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                                          it.d_image, it.d_beam, precision=precision, synchronize=True,
                                          w_proj=it.w_proj if it.w_proj.num_wplanes else None)
        single = b.images()
        # the oracle for two of them
        for k in (1, 4):
            it, s = b.items[k], b.snaps[k]
            oimg, obeam, _, _ = orc.image_visibilities(it.img_pars, s["vis"], s["weights"], s["uvw"],
                                                       wproj=it.w_proj if it.w_proj.num_wplanes else None)
            assert peak_rel_err(single[k][0], oimg) < (1e-10 if precision == capi.PRECISION_FP64 else 1e-4)
        # eager batch on 3 concurrent streams
        b.clear()
        totals = gpu.image_batch(b.items, precision=precision, max_concurrent=3)
        eager = b.images()
        for (a, ab), (c, cb) in zip(single, eager):
            assert peak_rel_err(c, a) <= tol and peak_rel_err(cb, ab) <= tol
        assert np.all(totals > 0)
        # graph: first call records, later calls replay; host work and syncs in between
        for rep in range(4):
            b.clear()
            gpu.image_batch(b.items, precision=precision, max_concurrent=3, flags=capi.BATCH_GRAPH)
            got = b.images()
            for (a, ab), (c, cb) in zip(eager, got):
                assert peak_rel_err(c, a) <= tol and peak_rel_err(cb, ab) <= tol, rep
        # new visibilities in the same buffers: the replay images the new data (nothing
        # of the first data is baked into the graph)
        for k, (N, cell, nv, ns, seed, wp, wmax) in enumerate(_specs()):
            sn = simdata.snapshot(N, cell, nv, ns + 2, seed=seed + 50, field_frac=0.5 if wp is not None else 0.25,
                                  w_max=wmax)
            gpu.htod(b.bufs[k][0], np.ascontiguousarray(sn["vis"]))
            gpu.htod(b.bufs[k][1], np.ascontiguousarray(sn["weights"]))
            gpu.htod(b.bufs[k][2], np.asfortranarray(sn["uvw"]))
        b.clear()
        gpu.image_batch(b.items, precision=precision, max_concurrent=3, flags=capi.BATCH_GRAPH)
        replayed = b.images()
        for it in b.items:
            gpu.image_visibilities_device(it.img_pars, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                                          it.d_image, it.d_beam, precision=precision, synchronize=True,
                                          w_proj=it.w_proj if it.w_proj.num_wplanes else None)
        fresh = b.images()
        for (a, ab), (c, cb), (o, _) in zip(fresh, replayed, eager):
            assert peak_rel_err(c, a) <= tol and peak_rel_err(cb, ab) <= tol
            assert peak_rel_err(a, o) > 1e-3  # the data did change
    finally:
        b.free()


def test_batch_graph_cache_is_bounded(gpu):
    """More distinct batches than the context keeps graphs for: the oldest graphs are
    dropped and re-recorded when their batch comes back, with the same images."""
    b = _Batch(gpu, [(256, 4.0, 5_000, 3, 711, None, 0.0)], capi.PRECISION_FP64)
    try:
        it = b.items[0]
        first = None
        for k in range(20):
            it.n_vis = 5_000 - k  # a new batch signature each call
            b.clear()
            gpu.image_batch(b.items, precision=capi.PRECISION_FP64, flags=capi.BATCH_GRAPH)
            if k == 0:
                first = b.images()[0][0]
        it.n_vis = 5_000
        for _ in range(2):
            b.clear()
            gpu.image_batch(b.items, precision=capi.PRECISION_FP64, flags=capi.BATCH_GRAPH)
            assert peak_rel_err(b.images()[0][0], first) <= 1e-12
    finally:
        b.free()


def test_autotune(gpu, orc, tmp_path):
    """Autotuning picks a tile width and an FFT split per shape, records them in the
    tune file, and the tuned configuration still matches the oracle."""
    tf = str(tmp_path / "tune.txt")
    gpu.set_autotune(1, tf)
    try:
        snap = simdata.workload_snapshot("c1")
        pars = capi.imager_pars(1024, 1.0)
        img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars)
        tcw, l1, us = gpu.get_tuning(pars, snap["vis"].size)
        assert tcw in (16, 32, 64) and us > 0
        assert open(tf).read().split()[1] == str(tcw)
        oimg, obeam, _, _ = orc.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
        assert peak_rel_err(img, oimg) < 1e-10 and peak_rel_err(beam, obeam) < 1e-10
        # every candidate configuration keeps the parity: force each through a tune file
        key = open(tf).read().split()[0]
        for t, split in ((16, 0), (32, 32), (64, 128)):
            forced = str(tmp_path / f"forced_{t}.txt")
            with open(forced, "w") as f:
                f.write(f"{key} {t} {split} 1.0\n")
            gpu.set_autotune(0, forced)
            assert gpu.get_tuning(pars, snap["vis"].size)[:2] == (t, split)
            img, beam = gpu.image_visibilities(snap["vis"], snap["weights"], snap["uvw"], pars)
            assert peak_rel_err(img, oimg) < 1e-10 and peak_rel_err(beam, obeam) < 1e-10, t
    finally:
        gpu.set_autotune(0, None)


# Fused batches: default kernels (7 taps), a 15-tap kernel (the wide tile class),
# no beam, padding 2 with gridding correction, a tiny snapshot, and two that take
# the streams beside them (W-projection, 17 taps).
FUSED_SPECS = [
    (512, 2.0, 20_000, 5, 721, None, 0.0),
    (1024, 1.0, 120_000, 12, 722, None, 0.0, {"kernel_support": 7, "oversampling": 8}),
    (256, 4.0, 5_000, 3, 723, None, 0.0, {"generate_beam": False}),
    (512, 2.0, 30_000, 8, 724, None, 0.0, {"padding_factor": 2.0, "gridding_correction": True}),
    (256, 20.0, 4_000, 6, 725, "w", 300.0),
    (256, 4.0, 6_000, 4, 726, None, 0.0, {"kernel_support": 8}),
    (2048, 1.0, 300_000, 40, 727, None, 0.0),
    (512, 2.0, 100, 1, 728, None, 0.0),
]
STREAMED = (4, 5)


def _fused_specs(order=None):
    sp = [(s[0], s[1], s[2], s[3], s[4], capi.wproj_pars(4, 8) if s[5] else None) + tuple(s[6:]) for s in FUSED_SPECS]
    return [sp[k] for k in order] if order is not None else sp


def _compare(items, got, want, tol):
    for k, (it, (a, ab), (c, cb)) in enumerate(zip(items, want, got)):
        assert peak_rel_err(c, a) <= tol, k
        if it.img_pars.generate_beam:
            assert peak_rel_err(cb, ab) <= tol, k


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32C, capi.PRECISION_FP32_TWOFLOAT])
def test_fused_batch_matches_single(gpu, orc, precision):
    """STP_BATCH_FUSED: the eligible snapshots are gridded by one launch per stage
    for each chunk (kernel timings count the launches) and give each snapshot's single-call image
    and beam (FP64 within 1e-12 of the peak, FP32 within 1e-6) with bit-identical
    sample_grid_totals; the streamed ones run beside them; a different batch next
    (other sizes and order) rebuilds the tables; graph mode replays the streamed part
    while the fused part runs eagerly; FP64 matches the oracle."""
    b = _Batch(gpu, _fused_specs(), precision)
    try:
        tol = 1e-12 if precision == capi.PRECISION_FP64 else 1e-6
        single, stotals = [], []
        for it in b.items:
            d = gpu.image_visibilities_device(it.img_pars, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                                              it.d_image, it.d_beam, precision=precision, synchronize=True,
                                              w_proj=it.w_proj if it.w_proj.num_wplanes else None)
            stotals.append(d.sample_grid_total)
        single = b.images()
        b.clear()
        gpu.set_timing(True)
        gpu.reset_timings()
        try:
            totals = gpu.image_batch(b.items, precision=precision, max_concurrent=2, flags=capi.BATCH_FUSED)
            t = gpu.timings()
        finally:
            gpu.set_timing(False)
        print({k: v for k, v in t.items() if k.startswith("batch_") or k.startswith("grid_")}, flush=True)
        # six fused snapshots: two chunks, one launch per stage each
        assert t["batch_count"][1] == 2 and t["batch_scatter"][1] == 2 and t["batch_total"][1] == 2
        assert 2 <= t["batch_tiles"][1] <= 4  # the 8-tap and the 16-tap class over the chunks
        assert t["grid_tiles"][1] == 1 and t["grid_wtiles"][1] >= 1  # 17 taps and W-projection on the streams
        _compare(b.items, b.images(), single, tol)
        assert list(totals) == stotals
        if precision == capi.PRECISION_FP64:
            for k in (1, 3):
                it, s = b.items[k], b.snaps[k]
                oimg, obeam, _, _ = orc.image_visibilities(it.img_pars, s["vis"], s["weights"], s["uvw"])
                assert peak_rel_err(single[k][0], oimg) < 1e-10 and peak_rel_err(single[k][1], obeam) < 1e-10
        # another batch: a subset in another order, then the first batch again
        order = [6, 3, 0, 7]
        sub = _Batch.__new__(_Batch)
        sub.gpu, sub.prec = gpu, precision
        sub.items = [b.items[k] for k in order]
        sub.snaps = [b.snaps[k] for k in order]
        sub.bufs = []
        sub.clear()
        totals = gpu.image_batch(sub.items, precision=precision, flags=capi.BATCH_FUSED)
        _compare(sub.items, sub.images(), [single[k] for k in order], tol)
        assert list(totals) == [stotals[k] for k in order]
        for rep in range(3):
            b.clear()
            totals = gpu.image_batch(b.items, precision=precision, max_concurrent=2,
                                     flags=capi.BATCH_FUSED | capi.BATCH_GRAPH)
            _compare(b.items, b.images(), single, tol)
            assert list(totals) == stotals, rep
    finally:
        b.free()


def test_fused_batch_only_eligible_and_autotuned(gpu, orc):
    """Every snapshot fused (no streams), with autotuned tile widths and FFT splits
    per shape: the FP32 images stay within 1e-4 of the oracle."""
    specs = [s for k, s in enumerate(_fused_specs()) if k not in STREAMED]
    b = _Batch(gpu, specs, capi.PRECISION_FP32C)
    try:
        gpu.image_batch(b.items, precision=capi.PRECISION_FP32C, flags=capi.BATCH_FUSED | capi.BATCH_AUTOTUNE)
        got = b.images()
        for k in (0, 1, 4):
            it, s = b.items[k], b.snaps[k]
            oimg, obeam, _, _ = orc.image_visibilities(it.img_pars, s["vis"], s["weights"], s["uvw"])
            assert peak_rel_err(got[k][0], oimg) < 1e-4 and peak_rel_err(got[k][1], obeam) < 1e-4, k
    finally:
        b.free()
