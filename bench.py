"""Benchmark of the STP imaging + source-finding hot path on one MI355X.

One step = one snapshot through stp_image_visibilities_device (gridding, c2r of
image and beam, normalisation) followed by stp_source_find_device, with the
visibilities already resident in HBM.  Workload c2: N=4096, cell 1", 1M
visibilities, 50 point sources (log-uniform 0.5-20 Jy), noise 0.1 Jy,
GaussianSinc support 3, oversampling 9, beam on; source finding at 5 sigma
detection / 3 sigma analysis, BINMEDIAN, 8-connectivity.

Three arithmetic paths run: the FP32 path with FP64 tile sums ("fp32c", the
headline `value`), the FP32 path with two-float tile sums ("fp32tf") and the
plain FP64 path ("fp64").  For each, a short
untimed pass brackets every launch with HIP events on the library's stream for
the per-kernel table; the timed region then carries events around the dominant
kernel only, whose average duration and algorithmic bytes per launch give
`roofline`.  The CPU baseline leg times the FP64 CPU oracle on the host and
reports each path's image error, and the source position and flux error,
against it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# LDS atomic-add rate of one MI355X, measured with every lane of 2048 workgroups of 512
# threads adding 1024 times to random addresses of a 32 x 71 LDS array (G atomics/s),
# by precision path: FP64 and FP32C sum with ds_add_f64 (1850 G/s); two-float with
# ds_add_f32 / ds_add_rtn_f32 (204 G/s, returning or not: FP32 LDS atomics run at a
# ninth of the FP64 rate on this part)
LDS_ATOMIC_PEAK_G = {0: 1850.0, 1: 1850.0, 2: 204.0}
METRIC = "snapshots/s + p50/p99 latency, FP32-comp vs FP64 GPU at <=1e-4 peak err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c1/c2: one snapshot per step (headline: c2); c3: W-projection imaging; "
                         "c4: 48-snapshot graph-replayed batches; c5: a served Poisson request stream")
    ap.add_argument("--snapshots", type=int, default=48, help="c4: snapshots per batch")
    ap.add_argument("--requests", type=int, default=60, help="c5: requests in the stream")
    ap.add_argument("--slots", type=int, default=4, help="c5: server slots")
    ap.add_argument("--precision", default="all", choices=["all", "both", "fp64", "fp32c", "fp32tf"],
                    help="paths to run (all: fp64, fp32c and fp32tf; both: fp64 and fp32c); "
                         "the headline is fp32c when it runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sourcefind", action="store_true")
    ap.add_argument("--cpu-snapshots", type=int, default=5, help="CPU baseline: snapshots timed after one warm-up")
    ap.add_argument("--streams", type=int, default=2,
                    help="c1/c2: also run the step on this many contexts at once, one host thread each "
                         "(snapshots overlap on the device); 1 = off")
    ap.add_argument("--no-c4", action="store_true", help="c2 run: skip the batched c4 sub-result")
    ap.add_argument("--no-c5", action="store_true", help="c2 run: skip the served c5 sub-result")
    return ap.parse_args()


def kernel_bytes(name, N, rows, nvis, n_entries, nb, S):
    """Algorithmic HBM bytes per launch (S = bytes per real of the grid and image).  With a
    beam (nb = 2) the beam grid is real, stored packed in half the bytes of the vis grid, and
    its FFT passes are launches of their own under the same names (u1, u2: the mean of the
    vis and beam launches)."""
    grid = rows * N * 2 * S  # one complex half-plane grid
    img = N * N * S
    grids = grid + (grid // 2 if nb == 2 else 0)
    if name in ("fft_u1", "fft_u2"):
        return 2 * grids // (2 if nb == 2 else 1)  # read + write every grid once, per launch
    if name == "fft_v_c2r":
        return grid + img               # read the vis grid, write the image
    if name == "fft_v_c2r_real":
        return grid // 2 + img          # read the packed beam grid, write the beam
    if name == "grid_tiles":
        return grids + 40 * n_entries   # write every grid element once, read the binned entries
    if name == "grid_count":
        return 48 * nvis                # read u, v, w(unused), vis, weight
    if name == "grid_scatter":
        return 48 * nvis + 40 * n_entries
    if name in ("sf_sum", "sel_hist1", "sel_hist2", "sel_gather", "sf_clip"):
        return img                      # one pass over the image
    if name == "ccl_tile":
        return img + 4 * N * N          # read the image, write the 32-bit pixel keys
    if name == "ccl_label_map":
        return 8 * N * N                # read the pixel keys, write the 32-bit label map
    return None


def pmc_traffic(workload, path, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (profiles/traffic_<workload>_<path>.txt: FETCH_SIZE + WRITE_SIZE, calibrated with
    a copy kernel), or None when no summary covers it."""
    fn = os.path.join(ROOT, "profiles", f"traffic_{workload}_{path}.txt")
    if not os.path.exists(fn):
        return None
    for line in open(fn):
        f = line.split()
        if line.startswith("#") or len(f) < 4:
            continue
        # the profile names the kernel function ("grid_tiles_os_kernel<float, AccF64, 8>"),
        # the library's timing names the launch ("grid_tiles")
        name = " ".join(f[:-3])
        if name == kernel or name.startswith(kernel + "_") or name.startswith(kernel + "<"):
            return int(f[-2]) + int(f[-1])
    return None


def measure(ctx, capi, args, pars, sfp, prec, n, N, bufs, do_sf):
    d_vis, d_w, d_uvw, d_img, d_beam, d_lab = bufs
    island_buf = ctx.island_buffer(1 << 16)  # page-locked: the catalogue arrives by DMA
    S = 8 if prec == capi.PRECISION_FP64 else 4
    rows = N // 2 + 1

    def step():
        ctx.image_visibilities_device(pars, n, d_vis, d_w, d_uvw, d_img, d_beam, precision=prec,
                                      synchronize=not do_sf)
        isl = None
        if do_sf:
            isl, _ = ctx.source_find_device(d_img, N, N, sfp, d_label_map=d_lab, precision=prec, out=island_buf)
        return isl

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    # gridding entries written and good visibilities, as the gridder reports them
    diag = ctx.image_visibilities_device(pars, n, d_vis, d_w, d_uvw, d_img, d_beam, precision=prec, synchronize=True)
    n_entries = int(diag.num_grid_entries)

    # per-kernel breakdown: every launch bracketed by HIP events (untimed pass)
    ctx.set_timing(True)
    ctx.set_timing_filter(None)
    ctx.reset_timings()
    nb_steps = max(1, min(5, args.steps))
    for _ in range(nb_steps):
        step()
    ctx.synchronize()
    kernels = {}
    for name, (ms, cnt) in ctx.timings().items():
        if name in ("image_visibilities", "fill"):
            continue
        avg_ms = ms / max(cnt, 1)
        kb = kernel_bytes(name, N, rows, n, n_entries, 2, S)
        kernels[name] = {"avg_us": round(avg_ms * 1000.0, 2), "per_step": round(cnt / nb_steps, 2),
                         "GBs": (round(kb / (avg_ms * 1e-3) / 1e9, 1) if kb else None)}
    dominant = max((k for k in kernels if kernel_bytes(k, N, rows, n, n_entries, 2, S)),
                   key=lambda k: kernels[k]["avg_us"] * kernels[k]["per_step"])

    # timed region: only the dominant kernel carries events (its live duration for the roofline)
    ctx.set_timing_filter([dominant])
    ctx.reset_timings()
    lat = []
    islands = None
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        islands = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - s0)
    ctx.synchronize()
    t_total = time.perf_counter() - t_start
    dom_ms, dom_cnt = ctx.timings()[dominant]
    ctx.set_timing(False)
    ctx.set_timing_filter(None)

    dom_bytes = kernel_bytes(dominant, N, rows, n, n_entries, 2, S)
    dom_avg_s = dom_ms * 1e-3 / max(dom_cnt, 1)
    achieved = dom_bytes / dom_avg_s / 1e9
    # LDS atomic adds of the tile kernel: 3 per tap (re, im, beam) and 49 taps per good
    # visibility (the conjugate copies near v = 0, under 1 %, are not counted)
    lds_atomics = None
    if dominant == "grid_tiles":
        per_launch = 3.0 * 49.0 * float(diag.num_good_vis) * (2 if prec == capi.PRECISION_FP32_TWOFLOAT else 1)
        rate = per_launch / dom_avg_s / 1e9
        lds_atomics = {"per_launch": per_launch, "achieved_G_per_s": round(rate, 1),
                       "op": "ds_add_rtn_f32 (hi) + ds_add_f32 (lo)" if prec == capi.PRECISION_FP32_TWOFLOAT
                       else "ds_add_f64",
                       "peak_G_per_s": LDS_ATOMIC_PEAK_G.get(prec),
                       "frac": (round(rate / LDS_ATOMIC_PEAK_G[prec], 4) if LDS_ATOMIC_PEAK_G.get(prec) else None)}
    ms_per_step = 1000.0 * t_total / args.steps
    # whole-snapshot compulsory bytes (beam on, 6 source-find passes over the image)
    B = 48 * n + 2 * (2 * rows * N * 2 * S) + 2 * N * N * S + (6 * N * N * S + 4 * N * N if do_sf else 0)
    image = np.empty((N, N), order="F", dtype=np.float64 if S == 8 else np.float32)
    ctx.dtoh(image, d_img)
    island_list = [dict(i) for i in islands] if islands is not None else None
    ctx.free_host(island_buf)
    return {
        "value": args.steps / t_total,
        "ms_per_step": ms_per_step,
        "p50_ms": float(np.percentile(lat, 50) * 1000.0),
        "p99_ms": float(np.percentile(lat, 99) * 1000.0),
        "roofline": {"bound": "hbm", "kernel": dominant, "avg_us": round(dom_avg_s * 1e6, 2),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "lds_atomics": lds_atomics,
        "path_roofline": {"compulsory_bytes": B, "achieved_GBs": round(B / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": round(B / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "islands": island_list,
        "image": image,
    }


def measure_pipelined(Context, capi, args, pars, sfp, prec, n, N, in_bufs, do_sf, nstreams):
    """The same step on `nstreams` contexts (one stream each, own outputs), each driven by
    its own host thread, so one snapshot's source finding overlaps the next one's imaging
    on the device.  Returns whole-job throughput and per-snapshot latencies."""
    import threading
    d_vis, d_w, d_uvw = in_bufs
    S = 8 if prec == capi.PRECISION_FP64 else 4
    ctxs, outs, isl_bufs = [], [], []
    for _ in range(nstreams):
        c = Context(0)
        ctxs.append(c)
        outs.append([c.malloc(S * N * N), c.malloc(S * N * N), c.malloc(4 * N * N)])
        isl_bufs.append(c.island_buffer(1 << 16))

    def step(k):
        c, (d_img, d_beam, d_lab) = ctxs[k], outs[k]
        c.image_visibilities_device(pars, n, d_vis, d_w, d_uvw, d_img, d_beam, precision=prec, synchronize=not do_sf)
        if do_sf:
            c.source_find_device(d_img, N, N, sfp, d_label_map=d_lab, precision=prec, out=isl_bufs[k])

    for k in range(nstreams):
        for _ in range(max(1, args.warmup)):
            step(k)
        ctxs[k].synchronize()
    lat = [[] for _ in range(nstreams)]
    share = [args.steps // nstreams + (1 if k < args.steps % nstreams else 0) for k in range(nstreams)]
    go = threading.Barrier(nstreams + 1)

    def worker(k):
        go.wait()
        for _ in range(share[k]):
            s0 = time.perf_counter()
            step(k)
            ctxs[k].synchronize()
            lat[k].append(time.perf_counter() - s0)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(nstreams)]
    for t in th:
        t.start()
    go.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    for k in range(nstreams):
        c = ctxs[k]
        for b in outs[k]:
            c.free(b)
        c.free_host(isl_bufs[k])
        c.close()
    all_lat = np.array([x for l in lat for x in l])
    return {"value": args.steps / wall, "ms_per_step": 1000.0 * wall / args.steps, "streams": nstreams,
            "p50_ms": float(np.percentile(all_lat, 50) * 1000.0), "p99_ms": float(np.percentile(all_lat, 99) * 1000.0)}


def bench_batch(args, emit=True):
    """Workload c4: 48 snapshots (N in {2048, 4096}, 2e4-2e6 visibilities) imaged per
    call of stp_image_batch, FP32 path with autotuned shapes, fused (one launch per
    stage across the snapshots) and as a replayed hipGraph of per-snapshot streams;
    compared with the FP64 path fused, per-snapshot streams run eagerly, and plain
    per-snapshot calls.  `value` is the faster FP32 form."""
    from fastimaging_amd import capi
    from tests import simdata
    from fastimaging_amd.stp import Context
    specs = simdata.batch_specs(args.snapshots)
    t0 = time.time()
    snaps = [simdata.snapshot(**sp) for sp in specs]
    gen_s = time.time() - t0
    ctx = Context(0)
    items, bufs = [], []
    for sp, sn in zip(specs, snaps):
        N, nv = sp["image_size"], sp["n_vis"]
        b = [ctx.malloc(16 * nv), ctx.malloc(8 * nv), ctx.malloc(24 * nv), ctx.malloc(8 * N * N), ctx.malloc(8 * N * N)]
        ctx.htod(b[0], np.ascontiguousarray(sn["vis"]))
        ctx.htod(b[1], np.ascontiguousarray(sn["weights"]))
        ctx.htod(b[2], np.asfortranarray(sn["uvw"]))
        it = capi.BatchItem()
        it.img_pars = capi.imager_pars(N, sp["cell_arcsec"])
        it.n_vis = nv
        it.d_vis, it.d_vis_weights, it.d_uvw_lambda, it.d_image, it.d_beam = [x.value for x in b]
        items.append(it)
        bufs.append(b)
    out = {}
    # the plain FP64 path: one stp_image_visibilities_device call per snapshot, in order,
    # no batch API, no autotuning, no graph
    for k, it in enumerate(items):
        ctx.image_visibilities_device(it.img_pars, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda, it.d_image,
                                      it.d_beam, precision=capi.PRECISION_FP64, synchronize=(k == len(items) - 1))
    reps = max(1, args.steps // 10)
    t1 = time.perf_counter()
    for _ in range(reps):
        for it in items:
            ctx.image_visibilities_device(it.img_pars, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                                          it.d_image, it.d_beam, precision=capi.PRECISION_FP64)
        ctx.synchronize()
    dt = (time.perf_counter() - t1) / reps
    out["fp64_plain"] = {"snapshots_per_s": round(len(items) / dt, 2), "ms_per_batch": round(dt * 1000, 3),
                         "batches_timed": reps}
    ref64 = {}
    check = sorted(range(len(items)), key=lambda k: (specs[k]["image_size"], specs[k]["n_vis"]))
    check = check[:3] + check[-3:]
    for k in check:
        N = specs[k]["image_size"]
        img = np.empty((N, N), order="F")
        ctx.dtoh(img, items[k].d_image)
        ref64[k] = img
    for name, prec, flags in (("fp32c_fused", capi.PRECISION_FP32C, capi.BATCH_FUSED | capi.BATCH_AUTOTUNE),
                              ("fp32c_graph", capi.PRECISION_FP32C, capi.BATCH_GRAPH | capi.BATCH_AUTOTUNE),
                              ("fp64_fused", capi.PRECISION_FP64, capi.BATCH_FUSED),
                              ("fp64_eager", capi.PRECISION_FP64, 0)):
        t1 = time.perf_counter()
        ctx.image_batch(items, precision=prec, flags=flags)  # tunes and records on the first call
        first = time.perf_counter() - t1
        ctx.image_batch(items, precision=prec, flags=flags)
        reps = max(1, args.steps // 10)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            ctx.image_batch(items, precision=prec, flags=flags)
        dt = (time.perf_counter() - t1) / reps
        out[name] = {"snapshots_per_s": round(len(items) / dt, 2), "ms_per_batch": round(dt * 1000, 3),
                     "first_call_s": round(first, 3), "batches_timed": reps}
    # accuracy of the FP32 batch: against the plain FP64 GPU images (themselves within
    # 1e-15 of the oracle in the tests) on the three smallest and three largest
    # snapshots, and against the FP64 oracle on the smallest
    acc = []
    ctx.image_batch(items, precision=capi.PRECISION_FP32C, flags=capi.BATCH_FUSED | capi.BATCH_AUTOTUNE)
    for k in check:
        N = specs[k]["image_size"]
        img = np.empty((N, N), order="F", dtype=np.float32)
        ctx.dtoh(img, items[k].d_image)
        a = {"snapshot": k, "N": N, "n_vis": specs[k]["n_vis"],
             "image_err_vs_fp64_gpu": float(np.max(np.abs(img - ref64[k])) / np.max(np.abs(ref64[k])))}
        if not args.no_cpu_baseline and k == check[0]:
            from oracle import oracle
            oimg, _, _, _ = oracle.image_visibilities(items[k].img_pars, snaps[k]["vis"], snaps[k]["weights"],
                                                      snaps[k]["uvw"])
            a["image_err_vs_oracle"] = float(np.max(np.abs(img - oimg)) / np.max(np.abs(oimg)))
        acc.append(a)
    best = max(("fp32c_fused", "fp32c_graph"), key=lambda k: out[k]["snapshots_per_s"])
    res = {"metric": "snapshots/s (48-snapshot batch)", "value": out[best]["snapshots_per_s"], "path": best,
           "unit": "snapshots/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic (deterministic SplitMix64 snapshots, inputs resident in HBM)",
           "config": {"workload": f"c4: {len(items)} snapshots, N in {{2048, 4096}}, 2e4-2e6 vis (total "
                                  f"{sum(sp['n_vis'] for sp in specs)}), GaussianSinc s3 os9, beam, imaging only",
                      "n_snapshots": len(items)},
           "paths": out,
           "speedup_fp32c_graph_vs_fp64_plain": round(out["fp32c_graph"]["snapshots_per_s"] /
                                                      out["fp64_plain"]["snapshots_per_s"], 3),
           "speedup_fp32c_graph_vs_fp64_eager": round(out["fp32c_graph"]["snapshots_per_s"] /
                                                      out["fp64_eager"]["snapshots_per_s"], 3),
           "speedup_fp32c_fused_vs_fp64_plain": round(out["fp32c_fused"]["snapshots_per_s"] /
                                                      out["fp64_plain"]["snapshots_per_s"], 3),
           "speedup_fp32c_fused_vs_fp64_fused": round(out["fp32c_fused"]["snapshots_per_s"] /
                                                      out["fp64_fused"]["snapshots_per_s"], 3),
           "accuracy": acc, "gen_s": round(gen_s, 2)}
    for b in bufs:
        for x in b:
            ctx.free(x)
    ctx.close()
    if emit:
        print(json.dumps(res), flush=True)
    return res


FP64_VALU_TFLOPS = 78.6  # MI355X FP64 vector peak (the W tile sums are FP64 FMAs on every path)


def bench_wproj(args):
    """Workload c3: W-projection imaging (N=4096, 2" cells, 2M visibilities with
    |w| <= 1500, 32 w-planes, support 32, oversampling 8, beam), FP64 and FP32
    paths.  The W tile kernel runs as several launches (phases) per snapshot; its
    time is their sum.  Roofline: 6 FP64 FMAs (12 flop) per tap against the FP64
    vector peak; beside it the W-kernel values read per tap (16 B FP64, 8 B FP32)
    against HBM, where the cache lives (L2 and MALL serve most of them)."""
    from fastimaging_amd import capi
    from tests import simdata
    from fastimaging_amd.stp import Context
    wl = simdata.WORKLOADS["c3"]
    t0 = time.time()
    snap = simdata.workload_snapshot("c3")
    gen_s = time.time() - t0
    N, n = wl["image_size"], wl["n_vis"]
    pars = capi.imager_pars(N, wl["cell_arcsec"], oversampling=8, generate_beam=True, analytic_gcf=False)
    wp = capi.wproj_pars(32, 32)
    ks = 2 * 32 + 1
    ctx = Context(0)
    bufs = [ctx.malloc(16 * n), ctx.malloc(8 * n), ctx.malloc(24 * n), ctx.malloc(8 * N * N), ctx.malloc(8 * N * N)]
    ctx.htod(bufs[0], np.ascontiguousarray(snap["vis"]))
    ctx.htod(bufs[1], np.ascontiguousarray(snap["weights"]))
    ctx.htod(bufs[2], np.asfortranarray(snap["uvw"]))
    paths = [("fp32c", capi.PRECISION_FP32C, 8.0), ("fp64", capi.PRECISION_FP64, 16.0)]
    steps = max(1, min(args.steps, 5))
    out = {}
    for name, prec, kv_bytes in paths:
        def step(sync=False):
            return ctx.image_visibilities_device(pars, n, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], precision=prec,
                                                 synchronize=sync, w_proj=wp)
        diag = step(True)
        ctx.set_timing(True)
        ctx.set_timing_filter(None)
        ctx.reset_timings()
        step(True)
        kernels = {k: {"avg_us": round(ms / c * 1000.0, 2), "per_step": c}
                   for k, (ms, c) in ctx.timings().items() if k not in ("image_visibilities", "fill")}
        ctx.set_timing_filter(["grid_wtiles"])
        ctx.reset_timings()
        ctx.synchronize()
        lat = []
        t1 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            step()
            ctx.synchronize()
            lat.append(time.perf_counter() - s0)
        wall = time.perf_counter() - t1
        gms, gcnt = ctx.timings()["grid_wtiles"]
        ctx.set_timing(False)
        ctx.set_timing_filter(None)
        g_s = gms * 1e-3 / steps  # all phases of one snapshot
        taps = float(ks * ks) * diag.num_good_vis
        flops = 12.0 * taps
        achieved = flops / g_s / 1e12
        peak = FP64_VALU_TFLOPS
        kbytes = kv_bytes * taps  # one complex W-kernel value read per tap
        kgbs = kbytes / g_s / 1e9
        out[name] = {"snapshots_per_s": round(steps / wall, 3), "ms_per_step": round(1000.0 * wall / steps, 3),
                     "p50_ms": round(float(np.percentile(lat, 50)) * 1000.0, 3),
                     "roofline": {"bound": "valu", "kernel": "grid_wtiles", "us_per_snapshot": round(g_s * 1e6, 1),
                                  "launches_per_snapshot": gcnt // steps,
                                  "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                                  "frac": round(achieved / peak, 4),
                                  "flops": flops, "note": "6 FP64 FMAs (12 flop) per tap, ks^2 taps per good "
                                                          "visibility (conjugate copies near v=0 not counted)"},
                     "kernel_reads": {"bytes": kbytes, "achieved_GBs": round(kgbs, 1), "peak": HBM_PEAK_GBS,
                                      "frac": round(kgbs / HBM_PEAK_GBS, 4),
                                      "note": f"{int(kv_bytes)} B of W-kernel cache read per tap; the cache "
                                              "lives in HBM, L2 and MALL serve most reads"},
                     "kernels": kernels, "num_good_vis": diag.num_good_vis}
    img = np.empty((N, N), order="F", dtype=np.float32)
    ctx.image_visibilities_device(pars, n, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4],
                                  precision=capi.PRECISION_FP32C, synchronize=True, w_proj=wp)
    ctx.dtoh(img, bufs[3])
    img64 = np.empty((N, N), order="F")
    ctx.image_visibilities_device(pars, n, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4],
                                  precision=capi.PRECISION_FP64, synchronize=True, w_proj=wp)
    ctx.dtoh(img64, bufs[3])
    res = {"metric": "W-projection snapshots/s (imaging)", "value": out["fp32c"]["snapshots_per_s"],
           "unit": "snapshots/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic (deterministic SplitMix64 point-source snapshot, inputs resident in HBM)",
           "config": {"workload": f"c3: N={N}, cell {wl['cell_arcsec']}\", {n} vis, |w| <= {wl['w_max']}, "
                                  f"{wl['n_sources']} sources, 32 w-planes, support 32, os 8, beam, imaging only"},
           "paths": out,
           "fp32c_vs_fp64_image_err": float(np.max(np.abs(img.astype(np.float64) - img64)) / np.max(np.abs(img64))),
           "gen_s": round(gen_s, 2)}
    if not args.no_cpu_baseline:
        from oracle import oracle
        t1 = time.perf_counter()
        oimg, _, _, _ = oracle.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"], wproj=wp)
        cpu_s = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": round(1.0 / cpu_s, 5), "unit": "snapshots/s", "cores": oracle.threads(),
                               "kind": "port", "sample": "1 full snapshot through the FP64 CPU oracle"}
        res["fp64_image_err_vs_oracle"] = float(np.max(np.abs(img64 - oimg)) / np.max(np.abs(oimg)))
    for b in bufs:
        ctx.free(b)
    ctx.close()
    print(json.dumps(res), flush=True)


def bench_serve(args, emit=True):
    """Workload c5: a served stream of imaging + source-finding requests (N in
    {1024..8192}, 1e4-4e6 visibilities, 30 % W-projection) arriving as a Poisson
    process at ~75 % of the server's measured capacity; reports latency
    percentiles (submit to result, host buffers in and out) and throughput."""
    import threading
    from fastimaging_amd import capi
    from tests import simdata
    from fastimaging_amd.serve import Server
    specs = simdata.serve_specs()
    t0 = time.time()
    pool = []
    for sp in specs:
        kw = {k: v for k, v in sp.items() if k != "wproj"}
        sn = simdata.snapshot(**kw)
        pars = capi.imager_pars(sp["image_size"], sp["cell_arcsec"])
        wp = capi.wproj_pars(32, 32) if sp["wproj"] else None
        pool.append((sn, pars, wp))
    gen_s = time.time() - t0
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    prec = capi.PRECISION_FP32C

    def submit(srv, k):
        sn, pars, wp = pool[k % len(pool)]
        return srv.submit(sn["vis"], sn["weights"], sn["uvw"], pars, w_proj=wp, precision=prec, sourcefind=sfp,
                          want_image=False, want_beam=False, island_cap=1 << 18)

    with Server(num_slots=args.slots) as srv:
        # warm every slot on every request shape, then measure capacity with a closed loop
        for r in [srv.wait(submit(srv, k)) for k in range(len(pool) * 2)]:
            pass
        t1 = time.perf_counter()
        ids = [submit(srv, k) for k in range(len(pool) * 2)]
        for i in ids:
            srv.wait(i)
        capacity = len(ids) / (time.perf_counter() - t1)
        lam = 0.75 * capacity
        rng = np.random.default_rng(20260)
        gaps = rng.exponential(1.0 / lam, args.requests)
        ids = [None] * args.requests

        def producer():
            t = time.perf_counter()
            for k in range(args.requests):
                t += gaps[k]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[k] = submit(srv, k)

        th = threading.Thread(target=producer)
        t1 = time.perf_counter()
        th.start()
        th.join()
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t1
    lat = np.array([(r["t_end"] - r["t_submit"]) * 1000.0 for r in res])
    svc = np.array([(r["t_end"] - r["t_start"]) * 1000.0 for r in res])
    out = {"metric": "served snapshots/s + p50/p99 latency (Poisson arrivals)", "value": round(len(res) / wall, 3),
           "unit": "snapshots/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "service_p50_ms": round(float(np.percentile(svc, 50)), 3),
           "arrival_rate_per_s": round(lam, 3), "capacity_per_s": round(capacity, 3),
           "data": "synthetic (deterministic SplitMix64 request pool, host buffers in and out)",
           "config": {"workload": f"c5: {len(pool)} request shapes cycled, N in {{1024..8192}}, 1e4-4e6 vis, 30 % "
                                  f"W-projection (32 planes, support 32), FP32 path, source finding, "
                                  f"{args.slots} slots", "requests": args.requests, "slots": args.slots},
           "gen_s": round(gen_s, 2)}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "c3":
        return bench_wproj(args)
    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c5":
        return bench_serve(args)

    from fastimaging_amd import capi
    from tests import simdata
    from fastimaging_amd.stp import Context

    wl = simdata.WORKLOADS[args.workload]
    t0 = time.time()
    snap = simdata.workload_snapshot(args.workload)
    gen_s = time.time() - t0
    N = wl["image_size"]
    n = wl["n_vis"]
    pars = capi.imager_pars(N, wl["cell_arcsec"], kernel_function=capi.KERNEL_GAUSSIANSINC, kernel_support=3,
                            kernel_exact=False, oversampling=9, generate_beam=True)
    sfp = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    do_sf = not args.no_sourcefind

    ctx = Context(0)
    bufs = [ctx.malloc(16 * n), ctx.malloc(8 * n), ctx.malloc(24 * n), ctx.malloc(8 * N * N), ctx.malloc(8 * N * N),
            ctx.malloc(4 * N * N)]
    ctx.htod(bufs[0], np.ascontiguousarray(snap["vis"]))
    ctx.htod(bufs[1], np.ascontiguousarray(snap["weights"]))
    ctx.htod(bufs[2], np.asfortranarray(snap["uvw"]))

    paths = {"all": ["fp64", "fp32c", "fp32tf"], "both": ["fp64", "fp32c"], "fp64": ["fp64"], "fp32c": ["fp32c"],
             "fp32tf": ["fp32tf"]}[args.precision]
    precs = {"fp64": capi.PRECISION_FP64, "fp32c": capi.PRECISION_FP32C, "fp32tf": capi.PRECISION_FP32_TWOFLOAT}
    res = {}
    for p in paths:
        res[p] = measure(ctx, capi, args, pars, sfp, precs[p], n, N, bufs, do_sf)
        if args.streams > 1:
            res[p]["pipelined"] = measure_pipelined(Context, capi, args, pars, sfp, precs[p], n, N, bufs[:3], do_sf,
                                                    args.streams)
    head_name = "fp32c" if "fp32c" in res else paths[0]
    head = res[head_name]

    for p, r in res.items():
        r["roofline"]["traffic"] = pmc_traffic(args.workload, p, r["roofline"]["kernel"])
    result = {
        "metric": METRIC,
        "value": round(head["value"], 3),
        "unit": "snapshots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "p50_ms": round(head["p50_ms"], 4),
        "p99_ms": round(head["p99_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "f32",
        "data": "synthetic (deterministic SplitMix64 point-source snapshot, inputs resident in HBM)",
        "config": {"workload": f"{args.workload}: N={N}, cell {wl['cell_arcsec']}\", {n} vis, "
                               f"{wl['n_sources']} sources, GaussianSinc s3 os9, beam, "
                               + ("source find 5/3 sigma BINMEDIAN 8-conn" if do_sf else "imaging only")
                               + {"fp32c": ", FP32 taps with FP64 tile sums, FP32 grids and FFT",
                                  "fp32tf": ", FP32 taps with two-float tile sums, FP32 grids and FFT",
                                  "fp64": ", FP64"}[head_name],
                   "image_size": N, "n_vis": n, "path": head_name},
        "pipelined": _rounded(head.get("pipelined")),
        "roofline": head["roofline"],
        "lds_atomics": head["lds_atomics"],
        "path_roofline": head["path_roofline"],
        "kernels": head["kernels"],
    }
    for p, r in res.items():
        if p != head_name:
            result[p] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                         if k in ("value", "ms_per_step", "p50_ms", "p99_ms", "roofline", "lds_atomics",
                                  "path_roofline", "kernels")}
            if r.get("pipelined"):
                result[p]["pipelined"] = _rounded(r["pipelined"])
    for p in ("fp32c", "fp32tf"):
        if "fp64" in res and p in res:
            result[f"speedup_{p}_vs_fp64"] = round(res[p]["value"] / res["fp64"]["value"], 3)
    result["islands"] = len(head["islands"]) if head["islands"] is not None else None

    if not args.no_cpu_baseline:
        from oracle import oracle
        nthr = oracle.threads()
        # one warm-up, then the median of k snapshots, imaging and source finding timed
        # apart (the stages of common_functions.cpp:172-205)
        k = max(1, args.cpu_snapshots)
        oimg = oisl = None
        t_img, t_sf = [], []
        for r in range(k + 1):
            t0 = time.perf_counter()
            oimg, _, _, _ = oracle.image_visibilities(pars, snap["vis"], snap["weights"], snap["uvw"])
            t1 = time.perf_counter()
            if do_sf:
                _, oisl, _ = oracle.source_find(oimg, sfp)
            t2 = time.perf_counter()
            if r:
                t_img.append(t1 - t0)
                t_sf.append(t2 - t1)
        tot = [a + b for a, b in zip(t_img, t_sf)]
        cpu_s = float(np.median(tot))
        result["cpu_baseline"] = {"value": round(1.0 / cpu_s, 4), "unit": "snapshots/s", "cores": nthr,
                                  "kind": "port", "cpu": _cpu_model(),
                                  "stages_ms": {"imager": round(1000 * float(np.median(t_img)), 1),
                                                "sourcefind": round(1000 * float(np.median(t_sf)), 1)},
                                  "sample": f"median of {k} full snapshots of the same workload (after one warm-up) "
                                            f"through the FP64 CPU oracle: threaded gridder and FFT, its own radix-2 "
                                            f"FFT (no FFTW here), threaded source finder"}
        peak = float(np.max(np.abs(oimg)))
        acc = {}
        for p, r in res.items():
            a = {"image_err": float(np.max(np.abs(r["image"].astype(np.float64) - oimg)) / peak)}
            if do_sf and r["islands"] is not None:
                key = lambda i: (i["sign"], i["extremum_x_idx"], i["extremum_y_idx"])
                same = [key(i) for i in r["islands"]] == [key(i) for i in oisl]
                a["islands_match_oracle"] = same
                a["n_islands"] = len(r["islands"])
                if same and oisl:
                    a["flux_err"] = max(abs(x["extremum_val"] - y["extremum_val"])
                                        for x, y in zip(r["islands"], oisl)) / peak
                    a["position_err_px"] = max(
                        float(np.hypot(x["moments_fit"][1] - y["moments_fit"][1],
                                       x["moments_fit"][2] - y["moments_fit"][2]))
                        for x, y in zip(r["islands"], oisl))
            acc[p] = a
        result["accuracy_vs_oracle"] = acc
    result["gen_s"] = round(gen_s, 2)
    for b in bufs:
        ctx.free(b)
    ctx.close()
    # the batched and served forms of the FP32 path on their own workloads
    if args.workload == "c2" and not args.no_c4:
        c4 = bench_batch(args, emit=False)
        result["c4"] = {k: c4[k] for k in ("value", "unit", "path", "config", "paths", "speedup_fp32c_graph_vs_fp64_plain",
                                           "speedup_fp32c_graph_vs_fp64_eager", "speedup_fp32c_fused_vs_fp64_plain",
                                           "speedup_fp32c_fused_vs_fp64_fused", "accuracy", "gen_s")}
    if args.workload == "c2" and not args.no_c5:
        c5 = bench_serve(args, emit=False)
        result["c5"] = {k: c5[k] for k in c5 if k not in ("metric", "n_gpus", "higher_is_better", "dtype", "data")}
    print(json.dumps(result), flush=True)


def _rounded(d):
    return None if d is None else {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
