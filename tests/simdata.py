"""Deterministic synthetic snapshots for the benchmark workloads and tests.

Random numbers come from SplitMix64 (counter-based, so any slice of a stream
can be regenerated) mapped to 53-bit uniforms; normals use Box-Muller.  The
seed of snapshot s of workload c is 0x57A90000 + 100*c + s.

Visibility recipe (u, v, w in wavelengths, l, m in radians):
    V = sum_k A_k exp(-2 pi i (u l_k + v m_k)) [exp(-2 pi i w (sqrt(1 - l_k^2 - m_k^2) - 1))]
        + complex Gaussian noise (sigma per component)
u, v are uniform in a disk of radius 0.45 / cell_rad, so every kernel centre
passes the gridder's bounds check.  With the imager's sign conventions a
source lands at un-shifted image column l/cell (mod N) and row m/cell (mod N).
"""
import math

import numpy as np

ARCSEC = math.pi / (180.0 * 3600.0)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def seed_for(config, snapshot=0):
    return 0x57A90000 + 100 * int(config) + int(snapshot)


class Stream:
    """SplitMix64 uniform stream on [0, 1)."""

    def __init__(self, seed):
        self.seed = np.uint64(seed)
        self.counter = 0

    def uniform(self, n):
        with np.errstate(over="ignore"):
            i = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
            z = self.seed + i * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.counter += n
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def snapshot(image_size, cell_arcsec, n_vis, n_sources, seed, amp_range=(1.0, 10.0), log_amp=False,
             field_frac=0.25, w_max=0.0, noise_sigma=0.1, chunk=1 << 18, reference_model=False):
    """Returns dict(uvw (n,3) F-order, vis complex128 (n,), weights (n,), sources list of (A, l, m)).

    Visibilities follow the benchmark recipe V = sum A exp(-2 pi i (u l + v m + w (n - 1))).
    reference_model=True uses instead the reference's own point-source model
    (src/stp/visibility/visibility.cpp:15-28), V = sum A n exp(+2 pi i (u l + v m + w (n - 1))),
    the sign convention its W-kernels (aw_projection.cpp:455-474) undo."""
    st = Stream(seed)
    cell = cell_arcsec * ARCSEC
    # sources
    if log_amp:
        a = np.exp(np.log(amp_range[0]) + st.uniform(n_sources) * (np.log(amp_range[1]) - np.log(amp_range[0])))
    else:
        a = amp_range[0] + st.uniform(n_sources) * (amp_range[1] - amp_range[0])
    half_field = image_size * cell * field_frac
    l = (2.0 * st.uniform(n_sources) - 1.0) * half_field
    m = (2.0 * st.uniform(n_sources) - 1.0) * half_field
    # uv disk
    radius = 0.45 / cell
    r = radius * np.sqrt(st.uniform(n_vis))
    phi = 2.0 * math.pi * st.uniform(n_vis)
    uvw = np.zeros((n_vis, 3), order="F")
    uvw[:, 0] = r * np.cos(phi)
    uvw[:, 1] = r * np.sin(phi)
    if w_max > 0.0:
        uvw[:, 2] = (2.0 * st.uniform(n_vis) - 1.0) * w_max
    vis = np.zeros(n_vis, dtype=np.complex128)
    nterm = np.sqrt(np.maximum(0.0, 1.0 - l * l - m * m)) - 1.0
    for b in range(0, n_vis, chunk):
        e = min(n_vis, b + chunk)
        u = uvw[b:e, 0][:, None]
        v = uvw[b:e, 1][:, None]
        ph = u * l[None, :] + v * m[None, :]
        if w_max > 0.0:
            ph = ph + uvw[b:e, 2][:, None] * nterm[None, :]
        if reference_model:
            vis[b:e] = ((a * (nterm + 1.0))[None, :] * np.exp(2j * math.pi * ph)).sum(axis=1)
        else:
            vis[b:e] = (a[None, :] * np.exp(-2j * math.pi * ph)).sum(axis=1)
    vis += noise_sigma * (st.normal(n_vis) + 1j * st.normal(n_vis))
    return dict(uvw=uvw, vis=vis, weights=np.ones(n_vis), sources=list(zip(a, l, m)),
                image_size=image_size, cell_arcsec=cell_arcsec)


# The benchmark workloads (sizes as quoted for the reference's headline metric).
WORKLOADS = {
    # N=1024, 1", 50k vis, 5 sources A~U[1,10] in the inner half, w=0
    "c1": dict(config=0, image_size=1024, cell_arcsec=1.0, n_vis=50_000, n_sources=5, amp_range=(1.0, 10.0),
               log_amp=False, field_frac=0.25, w_max=0.0),
    # N=4096, 1", 1M vis, 50 sources A log-uniform in [0.5, 20], w=0
    "c2": dict(config=1, image_size=4096, cell_arcsec=1.0, n_vis=1_000_000, n_sources=50, amp_range=(0.5, 20.0),
               log_amp=True, field_frac=0.25, w_max=0.0),
    # W-projection: N=4096, 2", 2M vis, |w| <= 1500, 40 sources over the full field
    "c3": dict(config=2, image_size=4096, cell_arcsec=2.0, n_vis=2_000_000, n_sources=40, amp_range=(1.0, 10.0),
               log_amp=False, field_frac=0.5, w_max=1500.0),
}


def workload_snapshot(name, snapshot_index=0, n_vis=None, image_size=None):
    wl = dict(WORKLOADS[name])
    cfg = wl.pop("config")
    if n_vis is not None:
        wl["n_vis"] = n_vis
    if image_size is not None:
        wl["image_size"] = image_size
    return snapshot(seed=seed_for(cfg, snapshot_index), **wl)


def batch_specs(n_snapshots=48, config=3):
    """Workload c4: per snapshot N in {2048, 4096}, cell 1", n_vis log-uniform in
    [2e4, 2e6], 1-40 point sources, w = 0.  Returns a list of snapshot() kwargs."""
    st = Stream(seed_for(config, 999))
    u = st.uniform(3 * n_snapshots).reshape(n_snapshots, 3)
    out = []
    for k in range(n_snapshots):
        N = 2048 if u[k, 0] < 0.5 else 4096
        nv = int(round(np.exp(np.log(2e4) + u[k, 1] * (np.log(2e6) - np.log(2e4)))))
        ns = 1 + int(u[k, 2] * 40) if u[k, 2] < 1.0 else 40
        out.append(dict(image_size=N, cell_arcsec=1.0, n_vis=nv, n_sources=min(ns, 40), seed=seed_for(config, k),
                        amp_range=(1.0, 10.0), log_amp=False, field_frac=0.25, w_max=0.0))
    return out


def serve_specs(n_requests=12, config=4, max_vis=4_000_000):
    """Workload c5: N in {1024, 2048, 4096, 8192}, n_vis log-uniform in [1e4, max_vis],
    30 % W-projection requests (|w| <= 1500, 32 planes, support <= 32, cell 2"),
    the others cell 1" and w = 0.  Returns snapshot() kwargs plus 'wproj' (bool)."""
    st = Stream(seed_for(config, 999))
    u = st.uniform(4 * n_requests).reshape(n_requests, 4)
    out = []
    for k in range(n_requests):
        N = (1024, 2048, 4096, 8192)[min(3, int(u[k, 0] * 4))]
        nv = int(round(np.exp(np.log(1e4) + u[k, 1] * (np.log(max_vis) - np.log(1e4)))))
        w = u[k, 2] < 0.3
        ns = 1 + min(39, int(u[k, 3] * 40))
        out.append(dict(image_size=N, cell_arcsec=2.0 if w else 1.0, n_vis=nv, n_sources=ns, seed=seed_for(config, k),
                        amp_range=(1.0, 10.0), log_amp=False, field_frac=0.5 if w else 0.25,
                        w_max=1500.0 if w else 0.0, wproj=w))
    return out

