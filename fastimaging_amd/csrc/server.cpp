// Serving loop with continuous batching (C-ABI stp_server_*, include/stp_capi.h).
//
// The reference images one snapshot per driver call (src/reduce/reduce.cpp:113-143,
// src/stp-python/stp_python.cpp:17-138).  Here a server keeps `num_slots` contexts
// on one device; each slot is a worker thread with its own HIP stream and its own
// resident kernel tables, FFT twiddles and grid/FFT/source-find workspace.  Jobs
// queue in arrival order and a free slot admits the head job once its estimated
// device footprint fits in the unreserved part of the HBM budget, so requests that
// arrive while others are on the device start as soon as a slot and memory free up,
// and their kernels overlap with the running jobs' kernels.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/stp_capi.h"
#include "gridder.h"

namespace {

using clock_type = std::chrono::steady_clock;

struct JobRec {
    stp_job job;
    uint64_t id = 0;
    uint64_t bytes = 0;
    stp_job_result res{};
    bool done = false;
    bool claimed = false; // a stp_server_wait has taken the result
};

// Per-slot device buffers, grown on demand and kept between jobs.
struct SlotBufs {
    void* in = nullptr;
    size_t in_bytes = 0;
    void* out = nullptr;
    size_t out_bytes = 0;
};

size_t round_up(size_t x) { return (x + 4095) & ~size_t(4095); }

int wplanes_of(const stp_job& j) { return int(j.w_proj.num_wplanes); }

} // namespace

struct stp_server {
    stp_server_cfg cfg{};
    uint64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<JobRec*> queue;
    std::map<uint64_t, std::unique_ptr<JobRec>> jobs;
    uint64_t next_id = 1;
    // admission ledger: the estimate of every running job plus the device memory each
    // slot keeps between jobs (its workspace, tables, twiddles and staging buffers)
    uint64_t reserved = 0;
    std::vector<uint64_t> resident;  // per slot
    bool release_idle = false;       // the queue head waits for memory idle slots hold
    uint32_t running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::string last_error;
    clock_type::time_point t0 = clock_type::now();

    double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
};

extern "C" uint64_t stp_job_device_bytes(const stp_job* j)
{
    if (!j)
        return 0;
    const stp_imager_pars& p = j->img_pars;
    const uint64_t N = p.padded_image_size, I = p.image_size, n = j->n_vis;
    const uint64_t S = j->precision == STP_PRECISION_FP64 ? 8 : 4;
    const uint64_t rows = N / 2 + 1;
    const bool wproj = wplanes_of(*j) > 0;
    const uint64_t s = wproj ? j->w_proj.max_wpconv_support : p.kernel_support;
    const uint64_t ks = 2 * s + 1;
    // gridder entry capacity at the narrowest tile width the autotuner may pick (the most
    // tiles per footprint)
    const uint64_t entries = stp_hip::gridder_entry_capacity(int(N), true, int(s), p.kernel_exact != 0, 16, size_t(n));
    uint64_t b = 0;
    b += 48 * n;                                // inputs
    b += 2 * rows * N * 2 * S;                  // vis + sampling grids
    b += stp_hip::GRID_ENTRY_BYTES * entries;   // binned entries
    b += 4 * 2 * ((rows + 63) / 64) * ((N + 31) / 32) * ((n + 4095) / 4096 + 1); // bin counts + offsets
    b += 2 * I * I * S + 4 * I * I;             // image, beam, label map
    // source-find workspace per pixel: parents 4 B; segment masks and offsets 0.3 B; one
    // component record per 16 pixels (tile regions) plus a pool of one per 256, each
    // with its record (72 B), island sums (64 B), label/keep words (8 B) and island
    // slot (96 B): 15.9 B; sigma-clipping candidates S B and their lane table 0.25 B.
    // Fixed: histograms (up to 256 blocks x 8194 bins), kernel tables and twiddles.
    b += 21 * I * I + S * I * I + (uint64_t(10) << 20);
    if (wproj) {
        const uint64_t os = std::max<uint32_t>(1, (p.oversampling >> 1) << 1);
        const uint64_t cs = (os / 2) * 2 + 1;
        const uint64_t P = j->w_proj.num_wplanes;
        b += 26 * n + 16 * n;                   // sort keys, indices, flags, scan, plane ids
        b += P * cs * cs * ks * (ks + 2 * stp_hip::W_KERNEL_PAD_COLS) * 16; // W-kernel caches (complex FP64, padded)
        if (S == 4)
            b += 2 * rows * N * 16;             // FP64 running sums between the W tile phases
        // work array (workarea * oversampling, gridder.h:467-486) for the W-kernel FFTs
        const uint32_t uo = j->w_proj.undersampling_opt;
        const uint64_t us = uo > 0 ? (uint64_t(1) << (uo - 1)) : 0;
        uint64_t workarea = 2;
        if (us > 0)
            while (workarea / us < ks)
                workarea *= 2;
        else
            while (workarea < N)
                workarea *= 2;
        const uint64_t arr = workarea * os;
        if (j->w_proj.hankel_opt) {
            // Hankel kernels: profile, transform, spline coefficients and half quadrant per
            // plane, and the DHT or DFT table
            const uint64_t n = j->w_proj.hankel_proj_slice ? arr : arr / 2;
            const uint64_t ko = (s + 1) * os;
            b += P * (2 * n * 16 + 10 * n * 8 + ko * (ko + 1) / 2 * 16) + (j->w_proj.hankel_proj_slice ? 16 * n : 8 * n * n);
        } else {
            // the image-domain and FFT arrays of one group of planes (at most 256 MB each)
            const uint64_t PB = std::max<uint64_t>(1, std::min<uint64_t>(P, (uint64_t(1) << 28) / (arr * arr * 16)));
            b += 2 * PB * arr * arr * 16;
        }
        b += 4 * P;
    }
    return b;
}

namespace {

void fail(JobRec* r, stp_status st, const std::string& msg)
{
    r->res.status = st;
    std::snprintf(r->res.error, sizeof(r->res.error), "%s", msg.c_str());
}

// Runs one job on a slot's context.  Returns the status (also stored in r->res).
void run_job(stp_server* s, stp_ctx* ctx, SlotBufs& b, JobRec* r)
{
    const stp_job& j = r->job;
    auto check = [&](stp_status st) {
        if (st != STP_OK) {
            fail(r, st, stp_ctx_last_error(ctx));
            return false;
        }
        return true;
    };
    const size_t n = j.n_vis;
    const size_t I = j.img_pars.image_size;
    const size_t osz = j.precision == STP_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    const size_t need_in = round_up(48 * n + 64);
    const size_t need_out = round_up(2 * osz * I * I) + round_up(4 * I * I);
    if (b.in_bytes < need_in) {
        if (b.in)
            stp_device_free(ctx, b.in);
        b.in = nullptr;
        b.in_bytes = 0;
        if (!check(stp_device_malloc(ctx, &b.in, need_in)))
            return;
        b.in_bytes = need_in;
    }
    if (b.out_bytes < need_out) {
        if (b.out)
            stp_device_free(ctx, b.out);
        b.out = nullptr;
        b.out_bytes = 0;
        if (!check(stp_device_malloc(ctx, &b.out, need_out)))
            return;
        b.out_bytes = need_out;
    }
    double* d_vis = (double*)b.in;
    double* d_w = d_vis + 2 * n;
    double* d_uvw = d_w + n;
    char* d_img = (char*)b.out;
    char* d_beam = d_img + osz * I * I;
    int32_t* d_lab = (int32_t*)((char*)b.out + round_up(2 * osz * I * I));
    if (n) {
        if (!check(stp_memcpy_htod(ctx, d_vis, j.vis, 16 * n)) || !check(stp_memcpy_htod(ctx, d_w, j.vis_weights, 8 * n)) ||
            !check(stp_memcpy_htod(ctx, d_uvw, j.uvw_lambda, 24 * n)))
            return;
    }
    const stp_wproj_pars* wp = j.w_proj.num_wplanes > 0 ? &j.w_proj : nullptr;
    const stp_kernel_spec* ks = j.use_kernel_spec ? &j.kernel : nullptr;
    if (!check(stp_image_visibilities_device(ctx, &j.img_pars, wp, nullptr, ks, n, d_vis, d_w, d_uvw, j.precision, d_img,
            j.img_pars.generate_beam ? d_beam : nullptr, 1, &r->res.diag)))
        return;
    const bool have_image = r->res.diag.image_rows > 0;
    auto copy_out = [&](double* dst, const char* src) {
        if (!dst || !have_image)
            return true;
        if (osz == sizeof(double))
            return check(stp_memcpy_dtoh(ctx, dst, src, sizeof(double) * I * I));
        std::vector<float> tmp(I * I);
        if (!check(stp_memcpy_dtoh(ctx, tmp.data(), src, sizeof(float) * I * I)))
            return false;
        for (size_t k = 0; k < I * I; k++)
            dst[k] = tmp[k];
        return true;
    };
    if (!copy_out(j.image, d_img))
        return;
    if (j.img_pars.generate_beam && !copy_out(j.beam, d_beam))
        return;
    if (j.source_find && have_image) {
        size_t ni = 0;
        const stp_status st = stp_source_find_device(ctx, d_img, j.precision, int32_t(I), int32_t(I), &j.sf_pars,
            j.label_map ? d_lab : nullptr, j.islands, j.island_cap, &ni, &r->res.stats);
        r->res.n_islands = ni;
        if (!check(st))
            return;
        if (j.label_map && !check(stp_memcpy_dtoh(ctx, j.label_map, d_lab, sizeof(int32_t) * I * I)))
            return;
    }
    (void)s;
    r->res.status = STP_OK;
}

// Device bytes a slot keeps between jobs.
uint64_t slot_bytes(stp_ctx* ctx, const SlotBufs& b)
{
    return (ctx ? stp_ctx_workspace_bytes(ctx) : 0) + b.in_bytes + b.out_bytes;
}

// Frees everything a slot keeps between jobs.
void slot_release(stp_ctx* ctx, SlotBufs& b)
{
    if (!ctx)
        return;
    stp_ctx_release_workspace(ctx);
    if (b.in)
        stp_device_free(ctx, b.in);
    if (b.out)
        stp_device_free(ctx, b.out);
    b = SlotBufs();
}

void worker(stp_server* s, int slot)
{
    stp_ctx* ctx = nullptr;
    const stp_status cst = stp_ctx_create(&ctx, s->cfg.device);
    SlotBufs bufs;
    for (;;) {
        JobRec* r = nullptr;
        bool admitted = false;
        bool released = false; // at most one release per idle period
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->release_idle && s->resident[slot] > 0 && !released) {
                    released = true;
                    // the queue head does not fit: give back what this idle slot holds
                    lk.unlock();
                    slot_release(ctx, bufs);
                    lk.lock();
                    s->reserved -= s->resident[slot];
                    s->resident[slot] = slot_bytes(ctx, bufs);
                    s->reserved += s->resident[slot];
                    s->cv_work.notify_all();
                    continue;
                }
                if (!s->queue.empty()) {
                    JobRec* head = s->queue.front();
                    if (head->bytes > s->budget || cst != STP_OK) {
                        // can never run: fail it without reserving
                        s->queue.pop_front();
                        r = head;
                        break;
                    }
                    if (s->reserved + head->bytes <= s->budget) {
                        s->queue.pop_front();
                        s->reserved += head->bytes;
                        s->running++;
                        s->release_idle = false;
                        r = head;
                        admitted = true;
                        break;
                    }
                    // it does not fit beside what is running and resident: this slot's
                    // own memory first, then every idle slot's (one notification when
                    // the request is raised; each release notifies again)
                    if (s->resident[slot] > 0 && !released) {
                        s->release_idle = true;
                        continue;
                    }
                    if (!s->release_idle) {
                        s->release_idle = true;
                        s->cv_work.notify_all();
                    }
                } else if (s->stop) {
                    break;
                } else {
                    s->release_idle = false;
                }
                s->cv_work.wait(lk);
            }
        }
        if (!r)
            break;
        r->res.slot = slot;
        r->res.t_start = s->now();
        r->res.job_bytes = r->bytes;
        if (cst != STP_OK)
            fail(r, cst, "stp_ctx_create failed for the server's device");
        else if (!admitted)
            fail(r, STP_ERR_OUT_OF_MEMORY, "job needs more device memory than the server's HBM budget");
        else {
            r->res.slot_bytes_before = slot_bytes(ctx, bufs);
            run_job(s, ctx, bufs, r);
            r->res.slot_bytes_after = slot_bytes(ctx, bufs);
        }
        r->res.t_end = s->now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (admitted) {
                // the job's estimate is replaced by what the slot now holds
                s->reserved -= r->bytes;
                s->reserved -= s->resident[slot];
                s->resident[slot] = r->res.slot_bytes_after;
                s->reserved += s->resident[slot];
                s->running--;
            }
            r->done = true;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    if (ctx) {
        slot_release(ctx, bufs);
        stp_ctx_destroy(ctx);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->reserved -= s->resident[slot];
    s->resident[slot] = 0;
}

} // namespace

extern "C" {

stp_status stp_server_create(stp_server** out, const stp_server_cfg* cfg)
{
    if (!out || !cfg || cfg->num_slots == 0 || cfg->num_slots > 64)
        return STP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return STP_ERR_DEVICE;
    if (cfg->device < 0 || cfg->device >= count)
        return STP_ERR_INVALID_ARGUMENT;
    stp_server* s = new (std::nothrow) stp_server();
    if (!s)
        return STP_ERR_OUT_OF_MEMORY;
    s->cfg = *cfg;
    s->resident.assign(cfg->num_slots, 0);
    s->budget = cfg->hbm_budget_bytes;
    if (s->budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(cfg->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return STP_ERR_DEVICE;
        }
        s->budget = uint64_t(double(fr) * 0.8);
    }
    try {
        for (uint32_t k = 0; k < cfg->num_slots; k++)
            s->workers.emplace_back(worker, s, int(k));
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto& t : s->workers)
            t.join();
        delete s;
        return STP_ERR_INTERNAL;
    }
    *out = s;
    return STP_OK;
}

stp_status stp_server_destroy(stp_server* s)
{
    if (!s)
        return STP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers)
        t.join();
    delete s;
    return STP_OK;
}

stp_status stp_server_submit(stp_server* s, const stp_job* job, uint64_t* job_id)
{
    if (!s || !job || !job_id)
        return STP_ERR_INVALID_ARGUMENT;
    if (job->n_vis > 0 && (!job->vis || !job->vis_weights || !job->uvw_lambda)) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = "job visibility buffers are NULL";
        return STP_ERR_INVALID_ARGUMENT;
    }
    if (job->source_find && job->island_cap > 0 && !job->islands) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = "job island buffer is NULL with a non-zero capacity";
        return STP_ERR_INVALID_ARGUMENT;
    }
    auto r = std::make_unique<JobRec>();
    r->job = *job;
    r->bytes = stp_job_device_bytes(job);
    r->res.status = STP_ERR_INTERNAL;
    r->res.slot = -1;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
            return STP_ERR_INVALID_ARGUMENT;
        r->id = s->next_id++;
        r->res.t_submit = s->now();
        *job_id = r->id;
        s->queue.push_back(r.get());
        s->jobs.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_all();
    return STP_OK;
}

stp_status stp_server_poll(stp_server* s, uint64_t job_id, int32_t* done)
{
    if (!s || !done)
        return STP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end())
        return STP_ERR_INVALID_ARGUMENT;
    *done = it->second->done ? 1 : 0;
    return STP_OK;
}

stp_status stp_server_wait(stp_server* s, uint64_t job_id, stp_job_result* result)
{
    if (!s)
        return STP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        s->last_error = "unknown job id (or already waited)";
        return STP_ERR_INVALID_ARGUMENT;
    }
    JobRec* r = it->second.get();
    if (r->claimed) {
        s->last_error = "job is already being waited for";
        return STP_ERR_INVALID_ARGUMENT;
    }
    r->claimed = true;
    s->cv_done.wait(lk, [&] { return r->done; });
    if (result)
        *result = r->res;
    s->jobs.erase(job_id);
    return STP_OK;
}

stp_status stp_server_load(stp_server* s, uint32_t* queued, uint32_t* running, uint64_t* reserved_bytes)
{
    if (!s)
        return STP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued)
        *queued = uint32_t(s->queue.size());
    if (running)
        *running = s->running;
    if (reserved_bytes)
        *reserved_bytes = s->reserved;
    return STP_OK;
}

const char* stp_server_last_error(const stp_server* s)
{
    return s ? s->last_error.c_str() : "NULL server";
}

} // extern "C"
