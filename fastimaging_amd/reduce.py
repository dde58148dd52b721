"""The reference's `reduce` driver (src/reduce/reduce.cpp:74-173) on the GPU library.

    python -m fastimaging_amd.reduce config.json data.npz [-d] [-s islands.json] [-o labelmap.npz]

Steps, as reduce.cpp: load the JSON configuration (load_json_config.h) and round the
image sizes up to multiples of 4 (common_functions.cpp:49-78); load uvw_lambda, vis
and snr_weights from the NPZ (load_data.h); with -d subtract the visibilities of the
NPZ's skymodel (visibility.cpp:15-40); image; source-find (with the least-squares
island fit when the configuration asks for it); write the islands as JSON
(save_json_sf_output.cpp:9-52) and the label map as NPZ.  Visibilities are uploaded
once and every stage runs on device buffers; the stage times are printed at the end
(common_functions.cpp:172-205).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import capi
from .config import ConfigurationFile, island_json, load_npy_complex_array, load_npy_double_array, set_image_sizes
from .stp import Context

PRECISIONS = {"fp64": capi.PRECISION_FP64, "fp32c": capi.PRECISION_FP32C, "fp32tf": capi.PRECISION_FP32_TWOFLOAT}


def run(config, npz, diff=False, out_json=None, out_npz=None, precision="fp64", ctx=None, log=print):
    """Runs the pipeline; returns dict(islands, fits, label_map, image, timings)."""
    t = {}
    t0 = time.perf_counter()
    cfg = config if isinstance(config, ConfigurationFile) else ConfigurationFile(config)
    for note in set_image_sizes(cfg.img_pars):
        log(f"warning: {note}")
    uvw = load_npy_double_array(npz, "uvw_lambda")
    vis = load_npy_complex_array(npz, "vis").reshape(-1)
    w = load_npy_double_array(npz, "snr_weights").reshape(-1)
    if uvw.shape[1] != 3 or uvw.shape[0] != vis.size or w.size != vis.size:
        raise ValueError("uvw_lambda must be n x 3 and vis, snr_weights of length n")
    sky = load_npy_double_array(npz, "skymodel") if diff else None
    t["load"] = time.perf_counter() - t0
    prec = PRECISIONS[precision]
    own = ctx is None
    ctx = ctx or Context(0)
    N = cfg.img_pars.image_size
    n = vis.size
    osz = 8 if prec == capi.PRECISION_FP64 else 4
    bufs = []
    try:
        t0 = time.perf_counter()
        d_vis, d_w, d_uvw = ctx.malloc(16 * n), ctx.malloc(8 * n), ctx.malloc(24 * n)
        bufs += [d_vis, d_w, d_uvw]
        ctx.htod(d_vis, np.ascontiguousarray(vis))
        ctx.htod(d_w, np.ascontiguousarray(w))
        ctx.htod(d_uvw, np.asfortranarray(uvw))
        if diff:
            # residual visibilities: input minus the skymodel's (reduce.cpp:113-122)
            ctx.model_visibilities_device(sky, n, d_uvw, d_vis, subtract=True)
        beam = bool(cfg.img_pars.generate_beam)
        d_img = ctx.malloc(osz * N * N)
        d_beam = ctx.malloc(osz * N * N) if beam else None
        d_lab = ctx.malloc(4 * N * N)
        bufs += [b for b in (d_img, d_beam, d_lab) if b is not None]
        ctx.synchronize()
        t["upload_residual"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        wp = cfg.w_proj if cfg.w_proj.num_wplanes > 0 else None
        diag = ctx.image_visibilities_device(cfg.img_pars, n, d_vis, d_w, d_uvw, d_img, d_beam, precision=prec,
                                             synchronize=True, w_proj=wp)
        t["imager"] = time.perf_counter() - t0
        if diag.image_rows == 0:
            raise RuntimeError("no visibility passed the bounds check: empty image")
        t0 = time.perf_counter()
        sfp = cfg.sourcefind_pars()
        islands, stats = ctx.source_find_device(d_img, N, N, sfp, d_label_map=d_lab, precision=prec,
                                                capacity=max(1 << 16, N * N // 256))
        islands = list(islands)
        t["sourcefind"] = time.perf_counter() - t0
        fits = None
        if cfg.gaussian_fitting:
            t0 = time.perf_counter()
            fits = ctx.fit_islands_device(d_img, N, N, d_lab, islands, precision=prec)
            t["gaussian_fitting"] = time.perf_counter() - t0
        image = np.empty((N, N), order="F", dtype=np.float64 if osz == 8 else np.float32)
        label_map = np.empty((N, N), order="F", dtype=np.int32)
        ctx.dtoh(image, d_img)
        ctx.dtoh(label_map, d_lab)
    finally:
        for b in bufs:
            ctx.free(b)
        if own:
            ctx.close()
    t0 = time.perf_counter()
    if out_json:
        with open(out_json, "w") as f:
            json.dump(island_json(islands, fits), f)
    if out_npz:
        np.savez(out_npz, label_map=label_map)
    t["save"] = time.perf_counter() - t0
    return dict(islands=islands, fits=fits, label_map=label_map, image=image, timings=t, stats=stats, config=cfg)


def main(argv=None):
    ap = argparse.ArgumentParser(description="Slow Transients Pipeline run on the GPU (reduce.cpp)")
    ap.add_argument("input_file_json", help="input JSON filename with configuration parameters")
    ap.add_argument("input_file_npz", help="input NPZ filename with simulation data (uvw_lambda, vis, snr_weights[, skymodel])")
    ap.add_argument("-s", "--output-file-json", help="output JSON filename for detected islands")
    ap.add_argument("-o", "--output-file-npz", help="output NPZ filename for the label map (label_map)")
    ap.add_argument("-d", "--diff", action="store_true",
                    help="use residual visibilities: input minus the visibilities of the NPZ's skymodel")
    ap.add_argument("-n", "--no-src", action="store_true", help="do not print the detected islands")
    ap.add_argument("-b", "--no-bench", action="store_true", help="do not print the stage timings")
    ap.add_argument("--precision", default="fp64", choices=sorted(PRECISIONS), help="imaging arithmetic (default fp64)")
    a = ap.parse_args(argv)
    r = run(a.input_file_json, a.input_file_npz, diff=a.diff, out_json=a.output_file_json, out_npz=a.output_file_npz,
            precision=a.precision)
    print(f"Number of detected sources: {len(r['islands'])}")
    if not a.no_src:
        for k, i in enumerate(r["islands"]):
            line = (f"Island {k}: label={i['label_idx']} sign={i['sign']} val={i['extremum_val']:.6g} "
                    f"x_idx={i['extremum_x_idx']} y_idx={i['extremum_y_idx']} samples={i['num_samples']} "
                    f"moments_fit={tuple(round(v, 6) for v in i['moments_fit'])}")
            if r["fits"] is not None:
                line += f" gaussian_fit={tuple(round(v, 6) for v in r['fits'][k]['leastsq_fit'])}"
            print(line)
    if not a.no_bench:
        for k, v in r["timings"].items():
            print(f"{k}: {v * 1000.0:.3f} ms")
    return 0


if __name__ == "__main__":
    sys.exit(main())
