"""The no-allocation imaging entry point (stp_image_visibilities_device_noalloc)
against the allocating one (stp_image_visibilities_device) through the C-ABI: two
calls in a row, the second on a smaller input, FP64 and FP32; a fresh context
holds no workspace after the no-allocation calls, and an input whose buffers do
not fit the scratch region is rejected with STP_ERR_CAPACITY."""
import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from fastimaging_amd.stp import Context, STPError
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu


def _upload(ctx, snap):
    nv = snap["weights"].size
    b = [ctx.malloc(16 * nv), ctx.malloc(8 * nv), ctx.malloc(24 * nv)]
    ctx.htod(b[0], np.ascontiguousarray(snap["vis"]))
    ctx.htod(b[1], np.ascontiguousarray(snap["weights"]))
    ctx.htod(b[2], np.asfortranarray(snap["uvw"]))
    return b, nv


def _image(ctx, pars, bufs, nv, precision, noalloc):
    N = pars.image_size
    osz, dt = (8, np.float64) if precision == capi.PRECISION_FP64 else (4, np.float32)
    d_img, d_beam = ctx.malloc(osz * N * N), ctx.malloc(osz * N * N)
    try:
        diag = ctx.image_visibilities_device(pars, nv, bufs[0], bufs[1], bufs[2], d_img, d_beam, precision=precision,
                                             synchronize=True, noalloc=noalloc)
        img = np.empty((N, N), order="F", dtype=dt)
        beam = np.empty((N, N), order="F", dtype=dt)
        ctx.dtoh(img, d_img)
        ctx.dtoh(beam, d_beam)
    finally:
        ctx.free(d_img)
        ctx.free(d_beam)
    return img.astype(np.float64), beam.astype(np.float64), diag


@pytest.mark.parametrize("precision,tol", [(capi.PRECISION_FP64, 1e-12), (capi.PRECISION_FP32C, 1e-6)])
def test_noalloc_matches_device_entry_point(precision, tol):
    # (N, cell, visibilities, sources, seed): the second call is smaller
    cases = [(1024, 1.0, 50_000, 5, 731), (512, 2.0, 20_000, 3, 732)]
    na = Context(0)   # only no-allocation calls
    ref = Context(0)  # the allocating entry point
    try:
        for N, cell, nv, ns, seed in cases:
            snap = simdata.snapshot(N, cell, nv, ns, seed=seed)
            pars = capi.imager_pars(N, cell)
            bufs, nv = _upload(na, snap)
            try:
                print(f"noalloc: N={N} n_vis={nv} precision={precision}", flush=True)
                img, beam, dg = _image(na, pars, bufs, nv, precision, noalloc=True)
                rimg, rbeam, rdg = _image(ref, pars, bufs, nv, precision, noalloc=False)
            finally:
                for b in bufs:
                    na.free(b)
            assert dg.num_good_vis == rdg.num_good_vis and dg.num_grid_entries == rdg.num_grid_entries
            assert dg.sample_grid_total == rdg.sample_grid_total
            e_i, e_b = peak_rel_err(img, rimg), peak_rel_err(beam, rbeam)
            print(f"noalloc: image err {e_i:.2e}, beam err {e_b:.2e}", flush=True)
            assert e_i <= tol and e_b <= tol
            assert np.abs(rimg).max() > 0
        # the no-allocation context never took a workspace buffer, table or twiddle table
        assert na.lib.stp_ctx_workspace_bytes(na.h) == 0
    finally:
        na.close()
        ref.close()


def test_noalloc_rejects_oversize():
    ctx = Context(0)
    try:
        # N = 4096 FP64 grids alone (2 x 2049 x 4096 x 16 B = 269 MB) exceed the region
        snap = simdata.snapshot(4096, 1.0, 1000, 1, seed=733)
        pars = capi.imager_pars(4096, 1.0)
        bufs, nv = _upload(ctx, snap)
        try:
            with pytest.raises(STPError) as ei:
                _image(ctx, pars, bufs, nv, capi.PRECISION_FP64, noalloc=True)
            assert ei.value.status == capi.STATUS_CODES["STP_ERR_CAPACITY"]
            # the context still works afterwards, on both entry points
            small = simdata.snapshot(256, 2.0, 5000, 2, seed=734)
            p2 = capi.imager_pars(256, 2.0)
            b2, n2 = _upload(ctx, small)
            try:
                a = _image(ctx, p2, b2, n2, capi.PRECISION_FP64, noalloc=True)
                b = _image(ctx, p2, b2, n2, capi.PRECISION_FP64, noalloc=False)
                assert peak_rel_err(a[0], b[0]) <= 1e-12
            finally:
                for x in b2:
                    ctx.free(x)
        finally:
            for b in bufs:
                ctx.free(b)
    finally:
        ctx.close()
