"""GPU tests of the serving loop (stp_server_*): every job's result is the same when
it runs alone as when it runs in a mix of concurrent jobs (FP64: within 1e-12 of the
peak, FP32 forms: within 1e-6; island sets and label maps identical), admission
keeps the reserved device memory within the HBM budget, and a job larger than the
budget fails cleanly."""
import threading
import time

import numpy as np
import pytest

from fastimaging_amd import capi
from tests import simdata
from fastimaging_amd.serve import Server
from fastimaging_amd.stp import STPError
from tests.helpers import peak_rel_err

pytestmark = pytest.mark.gpu


def _jobs():
    sf = capi.sourcefind_pars(5.0, 3.0, median_method=capi.MEDIAN_BINMEDIAN)
    out = []
    specs = [
        # (N, cell, n_vis, sources, seed, precision, wproj, w_max)
        (1024, 1.0, 50_000, 5, 901, capi.PRECISION_FP64, None, 0.0),
        (512, 2.0, 20_000, 8, 902, capi.PRECISION_FP32C, None, 0.0),
        (256, 20.0, 4_000, 6, 903, capi.PRECISION_FP64, capi.wproj_pars(4, 8), 300.0),
        (2048, 1.0, 200_000, 20, 904, capi.PRECISION_FP32C, None, 0.0),
        (512, 2.0, 30_000, 4, 905, capi.PRECISION_FP32_TWOFLOAT, None, 0.0),
        (1024, 1.0, 80_000, 10, 906, capi.PRECISION_FP64, None, 0.0),
    ]
    for N, cell, nv, ns, seed, prec, wp, wmax in specs:
        snap = simdata.snapshot(N, cell, nv, ns, seed=seed, field_frac=0.5 if wp is not None else 0.25, w_max=wmax)
        out.append(dict(snap=snap, pars=capi.imager_pars(N, cell), wp=wp, prec=prec, sf=sf))
    return out


def _submit(srv, j):
    s = j["snap"]
    return srv.submit(s["vis"], s["weights"], s["uvw"], j["pars"], w_proj=j["wp"], precision=j["prec"],
                      sourcefind=j["sf"], label_map=True)


def _same(a, b, prec):
    tol = 1e-12 if prec == capi.PRECISION_FP64 else 1e-6
    assert peak_rel_err(a["image"], b["image"]) <= tol
    assert peak_rel_err(a["beam"], b["beam"]) <= tol
    key = lambda i: (i["label_idx"], i["sign"], i["extremum_x_idx"], i["extremum_y_idx"], i["num_samples"])
    assert [key(i) for i in a["islands"]] == [key(i) for i in b["islands"]]
    assert np.array_equal(a["label_map"], b["label_map"])


def test_serve_alone_vs_mix(gpu, orc):
    jobs = _jobs()
    with Server(num_slots=1, hbm_budget_bytes=64 << 30) as srv:
        alone = [srv.wait(_submit(srv, j)) for j in jobs]
    # the admission estimate bounds what a job makes the slot hold
    assert alone[0]["slot_bytes_before"] == 0
    for a in alone:
        assert 0 < a["slot_bytes_after"] <= a["slot_bytes_before"] + a["job_bytes"]
        assert a["slot_bytes_after"] >= a["slot_bytes_before"]
    # one of them against the oracle: the served result is the imager's result
    j0 = jobs[0]
    oimg, obeam, _, _ = orc.image_visibilities(j0["pars"], j0["snap"]["vis"], j0["snap"]["weights"], j0["snap"]["uvw"])
    assert peak_rel_err(alone[0]["image"], oimg) < 1e-10
    # the mix: all jobs submitted at once from two threads onto 4 slots, so they overlap on the device
    with Server(num_slots=4, hbm_budget_bytes=64 << 30) as srv:
        ids = [None] * len(jobs)

        def sub(idx):
            for k in idx:
                ids[k] = _submit(srv, jobs[k])

        th = [threading.Thread(target=sub, args=(range(t, len(jobs), 2),)) for t in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mixed = [srv.wait(i) for i in ids]
    slots = {m["slot"] for m in mixed}
    assert len(slots) > 1  # the jobs ran on several slots
    overlap = any(a["t_start"] < b["t_end"] and b["t_start"] < a["t_end"]
                  for x, a in enumerate(mixed) for b in mixed[x + 1:])
    assert overlap
    for j, a, b in zip(jobs, alone, mixed):
        _same(a, b, j["prec"])


def test_serve_admission_budget(gpu):
    """A budget that fits one job at a time: jobs still all complete, the reserved bytes
    never exceed the budget, and a job bigger than the budget fails with OUT_OF_MEMORY."""
    all_jobs = _jobs()
    jobs = [all_jobs[0], all_jobs[1], all_jobs[4]]
    with Server(num_slots=3) as probe:
        need = [probe.job_bytes(j["pars"], j["snap"]["vis"].size, j["prec"], j["wp"]) for j in jobs]
    budget = max(need) + 1
    with Server(num_slots=3, hbm_budget_bytes=budget) as srv:
        ids = [_submit(srv, j) for j in jobs]
        seen = []
        t_end = time.time() + 60
        while not all(srv.poll(i) for i in ids) and time.time() < t_end:
            q, r, b = srv.load()
            seen.append((r, b))
            time.sleep(0.002)
        res = [srv.wait(i) for i in ids]
        assert all(b <= budget for _, b in seen)
        # resident slot memory is part of the ledger: what the slots keep after the
        # jobs is reserved, and never more than the budget
        q, r, b = srv.load()
        assert r == 0 and 0 < b <= budget
        for x in res:
            assert x["slot_bytes_after"] <= x["slot_bytes_before"] + x["job_bytes"]
        assert all(r["image"] is not None for r in res)
        # bigger than the budget: rejected when it reaches a slot
        big = all_jobs[3]
        jid = _submit(srv, big)
        with pytest.raises(STPError) as e:
            srv.wait(jid)
        assert e.value.status == capi.STATUS_CODES["STP_ERR_OUT_OF_MEMORY"]
