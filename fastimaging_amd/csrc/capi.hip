// extern "C" entry points of the gfx950 STP library (include/stp_capi.h).
// Every entry point converts exceptions into a status code and a message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>
#include <exception>

#include "common.h"
#include "fft.h"
#include "gridder.h"
#include "kernel_tables.h"
#include "reduce.h"
#include "sourcefind.h"
#include "visibility.h"

using namespace stp_hip;

namespace {

template <typename F>
stp_status guarded(stp_ctx* ctx, F f)
{
    if (!ctx)
        return STP_ERR_INVALID_ARGUMENT;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        f();
        ctx->last_error.clear();
        return STP_OK;
    } catch (const Error& e) {
        ctx->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host allocation failed";
        return STP_ERR_OUT_OF_MEMORY;
    } catch (const std::invalid_argument& e) {
        ctx->last_error = e.what();
        return STP_ERR_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return STP_ERR_INTERNAL;
    }
}

constexpr size_t MAX_BATCH_GRAPHS = 16; // captured batch graphs kept per context

bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }

// Parameter checks of image_visibilities (imager.h:257-287).
void check_imager_pars(const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_aproj_pars* ap)
{
    STP_REQUIRE(p != nullptr, "img_pars is NULL");
    STP_REQUIRE(p->padding_factor >= 1.0, "padding_factor must be >= 1");
    STP_REQUIRE(p->padded_image_size >= p->image_size, "padded_image_size < image_size");
    STP_REQUIRE(p->padded_image_size > 0 && is_pow2(p->padded_image_size), "padded image size must be a power of two");
    STP_REQUIRE(p->image_size >= 4 && p->image_size % 4 == 0, "image_size must be a multiple of 4");
    STP_REQUIRE(p->kernel_support > 0, "kernel_support must be > 0");
    STP_REQUIRE(p->cell_size > 0.0, "cell_size must be > 0");
    STP_REQUIRE(p->kernel_exact || p->oversampling >= 1, "oversampling must be >= 1 when kernel_exact is false");
    STP_REQUIRE(p->kernel_function >= STP_KERNEL_TOPHAT && p->kernel_function <= STP_KERNEL_PSWF, "unknown kernel function");
    if (ap && ap->num_timesteps > 0)
        throw Error(STP_ERR_UNSUPPORTED, "A-projection is not supported");
    if (wp && wp->num_wplanes > 0) {
        STP_REQUIRE(!p->kernel_exact, "W-projection requires kernel_exact == false (imager.h:287)");
        STP_REQUIRE(wp->interp_type >= 0 && wp->interp_type <= 2, "interp_type must be 0 (linear), 1 (cosine) or 2 (cubic)");
    }
}

struct ImagingPlan {
    int N;        // padded size
    int rows;     // N/2 + 1
    size_t grid_elems;
};

// Full imaging pipeline on device buffers: grid -> c2r (image, beam) -> normalise.
// The functor the ImageVisibilities constructor builds from the parameters
// (imager.cpp:25-61): trunc / half_base_width = kernel_support, default widths.
stp_kernel_spec dispatch_kernel(const stp_imager_pars* p)
{
    stp_kernel_spec k;
    stp_kernel_spec_init(&k, p->kernel_function, double(p->kernel_support));
    return k;
}

void check_kernel_spec(const stp_kernel_spec* k)
{
    STP_REQUIRE(k->kernel_function >= STP_KERNEL_TOPHAT && k->kernel_function <= STP_KERNEL_PSWF, "unknown kernel function");
    STP_REQUIRE(std::isfinite(k->param) && std::isfinite(k->width_normalization_gaussian) &&
            std::isfinite(k->width_normalization_sinc) && k->width_normalization_gaussian != 0.0 &&
            k->width_normalization_sinc != 0.0,
        "kernel parameters must be finite and widths non-zero");
    if (k->kernel_function == STP_KERNEL_TOPHAT || k->kernel_function == STP_KERNEL_TRIANGLE ||
        k->kernel_function == STP_KERNEL_PSWF)
        STP_REQUIRE(k->param > 0.0, "half_base_width / trunc must be > 0");
}

// Imaging shape of the autotuner: padded size, support, oversampling, kernel
// mode, precision, W-projection planes and the visibility count's octave.
std::string tune_key(const stp_imager_pars* p, const stp_wproj_pars* wp, size_t n, int precision)
{
    int oct = 0;
    while ((size_t(1) << oct) <= n)
        oct++;
    char k[160];
    snprintf(k, sizeof(k), "N%u:s%u:os%u:x%d:p%d:w%u:ws%u:n%d", p->padded_image_size, p->kernel_support, p->oversampling,
        p->kernel_exact, precision, wp ? wp->num_wplanes : 0u, wp ? wp->max_wpconv_support : 0u, oct);
    return k;
}

// Gridder arguments of one imaging call (imager.h:233-376) with the autotuned
// launch shape of its imaging shape; the FFT's u split in *fft_l1.
GridderArgs imager_args(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_kernel_spec& kern,
    size_t n, int precision, int* fft_l1)
{
    const int N = int(p->padded_image_size);
    GridderArgs a;
    a.image_size = N;
    a.support = int(p->kernel_support);
    a.oversampling = int(p->oversampling);
    a.kernel_exact = p->kernel_exact != 0;
    a.shift_uv = true;
    a.halfplane = true;
    a.generate_beam = p->generate_beam != 0;
    a.uv_scale = ((p->cell_size / 3600.0) * (M_PI / 180.0)) * double(N); // imager.h:293
    a.kernel = kern;
    a.twofloat = precision == STP_PRECISION_FP32_TWOFLOAT;
    if (wp && wp->num_wplanes > 0)
        a.wproj = *wp;
    a.cell_rad = (p->cell_size / 3600.0) * (M_PI / 180.0);
    a.analytic_gcf = p->analytic_gcf != 0;
    *fft_l1 = 0;
    auto tc = ctx->tune.find(tune_key(p, wp, n, precision));
    if (tc != ctx->tune.end()) {
        a.tile_cols = tc->second.tile_cols;
        *fft_l1 = tc->second.fft_l1;
    }
    return a;
}

// Gridding correction table of the padded size (imager.h:96-213), or NULL when
// the call has none.
const double* gcf_table(stp_ctx* ctx, const stp_imager_pars* p, const stp_kernel_spec& kern)
{
    if (!p->gridding_correction)
        return nullptr;
    const int N = int(p->padded_image_size);
    const KernelFn kf = kernel_fn(kern);
    char key[192];
    snprintf(key, sizeof(key), "gcf:%d:%.17g:%.17g:%.17g:%d:%d", kf.kind, kf.param, kf.width_g, kf.width_s, N, p->analytic_gcf);
    return has_table(ctx, key) ? dev_table(ctx, key, {})
                               : dev_table(ctx, key, img_dom_kernel(kf, size_t(N), false, p->analytic_gcf != 0));
}

template <typename T, typename OT>
void run_imager(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_kernel_spec* kspec, size_t n,
    const double* d_vis, const double* d_w, const double* d_uvw, OT* d_image, OT* d_beam, double* d_total,
    unsigned long long* d_ngood, bool twofloat, unsigned long long* d_nentries = nullptr)
{
    const stp_kernel_spec kern = kspec ? *kspec : dispatch_kernel(p);
    check_kernel_spec(&kern);
    const int N = int(p->padded_image_size);
    const bool beam = p->generate_beam != 0;
    STP_REQUIRE(!beam || d_beam != nullptr, "beam buffer is NULL with generate_beam");
    int fft_l1 = 0;
    const int precision = std::is_same<T, double>::value ? STP_PRECISION_FP64
                                                          : (twofloat ? STP_PRECISION_FP32_TWOFLOAT : STP_PRECISION_FP32C);
    GridderArgs a = imager_args(ctx, p, wp, kern, n, precision, &fft_l1);
    a.d_nentries = d_nentries;
    // without W-projection the sampling grid is real: packed, half the bytes in the
    // grid and the FFT
    a.beam_packed = beam && a.wproj.num_wplanes == 0 && N >= 8;
    const int rows = N / 2 + 1;
    const size_t ge = size_t(rows) * N, gb = a.beam_packed ? ge / N * (N / 2) : ge;
    cplx<T>* grids = (cplx<T>*)ws_get(ctx, "imager_grids", sizeof(cplx<T>) * (ge + (beam ? gb : 0)));
    grid_visibilities<T>(ctx, a, n, d_uvw, d_vis, d_w, grids, beam ? grids + ge : nullptr, d_total, d_ngood);
    NormSpec ns;
    ns.image_size = int(p->image_size);
    ns.total_ptr = d_total;
    ns.u_l1 = fft_l1;
    ns.gcf = gcf_table(ctx, p, kern);
    OutPtrs<OT> outs{};
    outs.p[0] = d_image;
    if (a.beam_packed) {
        fft_c2r_2d<T, OT>(ctx, grids, 1, N, N, outs, ns);
        OutPtrs<OT> bouts{};
        bouts.p[0] = d_beam;
        fft_c2r_2d_real<T, OT>(ctx, grids + ge, 1, N, bouts, ns);
    } else {
        outs.p[1] = d_beam;
        fft_c2r_2d<T, OT>(ctx, grids, beam ? 2 : 1, N, N, outs, ns);
    }
}

// Whether a batch item can join the fused launches (see grid_batch_eligible).
bool batch_fusable(stp_ctx* ctx, const stp_batch_item& it, const stp_kernel_spec* kspec, int precision)
{
    if (it.w_proj.num_wplanes > 0)
        return false;
    const stp_kernel_spec kern = kspec ? *kspec : dispatch_kernel(&it.img_pars);
    int l1;
    return grid_batch_eligible(imager_args(ctx, &it.img_pars, nullptr, kern, it.n_vis, precision, &l1), it.n_vis);
}

// A chunk of the fused items of a batch: one gridding launch per stage over all of
// them (grid_batch_bin, then grid_batch_tiles), then one c2r per group of equal
// transform shape (padded size, image size, gcf, u split, beam) over the group's
// grids, which lie contiguous.  The three phases run on the context stream current
// at each call, so a caller can give them different, ordered streams.
template <typename T, typename OT>
struct FusedChunk {
    struct Shape {
        int N, image, l1, beam;
        const double* gcf;
        bool operator<(const Shape& o) const
        {
            return std::tie(N, image, l1, beam, gcf) < std::tie(o.N, o.image, o.l1, o.beam, o.gcf);
        }
    };
    struct Run {
        Shape s;
        cplx<T>* grids;
        int nbatch;
        bool real;       // packed real beam grids
        size_t out_off, tot_off;
    };
    std::vector<Run> runs;
    BatchPlan plan;

    void bin(stp_ctx* ctx, const stp_batch_item* items, const std::vector<size_t>& idx, const stp_kernel_spec* kspec,
        int precision, double* totals)
    {
        std::vector<BatchGridItem> bg(idx.size());
        std::map<Shape, std::vector<size_t>> groups; // shape -> positions in bg
        for (size_t q = 0; q < idx.size(); q++) {
            const stp_batch_item& it = items[idx[q]];
            const stp_imager_pars* p = &it.img_pars;
            const stp_kernel_spec kern = kspec ? *kspec : dispatch_kernel(p);
            check_kernel_spec(&kern);
            int l1 = 0;
            BatchGridItem& b = bg[q];
            b.a = imager_args(ctx, p, nullptr, kern, it.n_vis, precision, &l1);
            b.a.beam_packed = p->generate_beam != 0;  // eligible snapshots: no W-projection, N >= 8
            b.n = it.n_vis;
            b.uv = it.d_uvw_lambda;
            b.vis = it.d_vis;
            b.wts = it.d_vis_weights;
            b.total = totals + 2 * idx[q];
            b.ngood = (unsigned long long*)(totals + 2 * idx[q] + 1);
            groups[Shape{ int(p->padded_image_size), int(p->image_size), l1, p->generate_beam ? 1 : 0,
                       gcf_table(ctx, p, kern) }]
                .push_back(q);
        }
        // grids: each group's vis grids contiguous in item order, then its packed beam grids
        size_t elems = 0;
        for (auto& gr : groups) {
            const size_t rows = size_t(gr.first.N / 2 + 1);
            elems += rows * (gr.first.N + (gr.first.beam ? gr.first.N / 2 : 0)) * gr.second.size();
        }
        cplx<T>* grids = (cplx<T>*)ws_get(ctx, "bat_grids", sizeof(cplx<T>) * std::max<size_t>(elems, 1));
        // per run: output and total pointer per grid, in the uploaded table
        std::vector<uint8_t> extra;
        runs.clear();
        size_t cur = 0;
        auto add_run = [&](const Shape& s, bool real, const std::vector<size_t>& members) {
            const size_t ge = size_t(s.N / 2 + 1) * (real ? s.N / 2 : s.N);
            Run r{ s, grids + cur, int(members.size()), real, 0, 0 };
            std::vector<OT*> outs;
            std::vector<const double*> tots;
            for (size_t q : members) {
                const stp_batch_item& it = items[idx[q]];
                (real ? bg[q].bgrid : bg[q].vgrid) = grids + cur;
                cur += ge;
                outs.push_back((OT*)(real ? it.d_beam : it.d_image));
                tots.push_back(bg[q].total);
            }
            r.out_off = extra.size();
            extra.insert(extra.end(), (const uint8_t*)outs.data(), (const uint8_t*)(outs.data() + outs.size()));
            r.tot_off = extra.size();
            extra.insert(extra.end(), (const uint8_t*)tots.data(), (const uint8_t*)(tots.data() + tots.size()));
            runs.push_back(r);
        };
        for (auto& gr : groups) {
            add_run(gr.first, false, gr.second);
            if (gr.first.beam)
                add_run(gr.first, true, gr.second);
        }
        plan = grid_batch_bin<T>(ctx, bg, extra);
    }

    void tiles(stp_ctx* ctx) { grid_batch_tiles<T>(ctx, plan); }

    void fft(stp_ctx* ctx)
    {
        const uint8_t* d_extra = (const uint8_t*)plan.d_extra;
        for (const Run& r : runs) {
            NormSpec ns;
            ns.image_size = r.s.image;
            ns.total_tab = (const double* const*)(d_extra + r.tot_off);
            ns.u_l1 = r.s.l1;
            ns.gcf = r.s.gcf;
            OutPtrs<OT> outs{};
            outs.tab = (OT* const*)(d_extra + r.out_off);
            if (r.real)
                fft_c2r_2d_real<T, OT>(ctx, r.grids, r.nbatch, r.s.N, outs, ns);
            else
                fft_c2r_2d<T, OT>(ctx, r.grids, r.nbatch, r.s.N, r.s.N, outs, ns);
        }
    }
};

// Fused items in chunks on two streams: every chunk's binning, then every chunk's
// tile sums, in turn on the first stream, and each chunk's FFTs on the second once
// its tile sums are done, so the tile sums of chunk c (LDS-bound) overlap the FFTs
// of chunk c-1 (HBM-bound), while the binning, whose 1024-thread workgroups wait
// long for whole CUs beside other kernels, runs alone.  (Two streams, not three:
// the device's few hardware queues may put a third on the queue of one of them.)
// events[c] marks the end of chunk c's tile sums.
template <typename T, typename OT>
void run_fused_chunks(stp_ctx* ctx, const stp_batch_item* items, const std::vector<std::vector<size_t>>& part,
    const stp_kernel_spec* kspec, int precision, double* totals, const hipStream_t* streams, const hipEvent_t* events,
    int slot0)
{
    const int P = int(part.size());
    std::vector<FusedChunk<T, OT>> ch(P);
    ctx->stream = streams[0];
    for (int c = 0; c < P; c++) {
        ctx->ws_slot = slot0 + c;
        ch[c].bin(ctx, items, part[c], kspec, precision, totals);
    }
    for (int c = 0; c < P; c++) {
        ctx->stream = streams[0];
        ctx->ws_slot = slot0 + c;
        ch[c].tiles(ctx);
        HIP_CHECK(hipEventRecord(events[c], streams[0]));
        HIP_CHECK(hipStreamWaitEvent(streams[1], events[c], 0));
        ctx->stream = streams[1];
        ch[c].fft(ctx);
    }
}

// Imaging on device buffers in the requested precision.
void run_imaging(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_kernel_spec* kernel, size_t n,
    const double* d_vis, const double* d_w, const double* d_uvw, int32_t precision, void* d_image, void* d_beam,
    double* d_total, unsigned long long* d_ngood, unsigned long long* d_nentries = nullptr)
{
    STP_REQUIRE(precision >= STP_PRECISION_FP64 && precision <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
    if (precision == STP_PRECISION_FP64)
        run_imager<double, double>(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, (double*)d_image, (double*)d_beam, d_total, d_ngood,
            false, d_nentries);
    else
        run_imager<float, float>(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, (float*)d_image, (float*)d_beam, d_total, d_ngood,
            precision == STP_PRECISION_FP32_TWOFLOAT, d_nentries);
}

// Autotuner: times the imaging of this very input with each candidate launch
// configuration (gridder tile width 16 / 32 / 64 columns, then the u-axis FFT
// split around the balanced one) and keeps the fastest for the shape.
void autotune_shape(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_kernel_spec* kernel, size_t n,
    const double* d_vis, const double* d_w, const double* d_uvw, int32_t precision, void* d_image, void* d_beam,
    double* d_total, unsigned long long* d_ngood)
{
    const std::string key = tune_key(p, wp, n, precision);
    if (ctx->tune.count(key))
        return;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const bool timing = ctx->timing;
    ctx->timing = false;
    auto time_one = [&](const TuneChoice& c) {
        ctx->tune[key] = c;
        run_imaging(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, precision, d_image, d_beam, d_total, d_ngood); // warm
        HIP_CHECK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 2; r++)
            run_imaging(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, precision, d_image, d_beam, d_total, d_ngood);
        HIP_CHECK(hipEventRecord(e1, ctx->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return double(ms) * 500.0; // us per run
    };
    TuneChoice best;
    best.us = 1e300;
    try {
        for (int tcw : { 16, 32, 64 }) {
            TuneChoice c;
            c.tile_cols = tcw;
            c.us = time_one(c);
            if (c.us < best.us)
                best = c;
        }
        const int N = int(p->padded_image_size);
        int lg = 0;
        while ((1 << lg) < N)
            lg++;
        const int bal = 1 << ((lg + 1) / 2);
        for (int l1 : { bal / 2, bal * 2 }) {
            if (l1 < 2 || l1 > 128 || N / l1 < 2 || N / l1 > 128)
                continue;
            TuneChoice c = best;
            c.fft_l1 = l1;
            c.us = time_one(c);
            if (c.us < best.us)
                best = c;
        }
    } catch (...) {
        ctx->tune.erase(key);
        ctx->timing = timing;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    ctx->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tune[key] = best;
    if (!ctx->tune_file.empty()) {
        if (FILE* f = std::fopen(ctx->tune_file.c_str(), "a")) {
            std::fprintf(f, "%s %d %d %.3f\n", key.c_str(), best.tile_cols, best.fft_l1, best.us);
            std::fclose(f);
        }
    }
}

void load_tune_file(stp_ctx* ctx, const std::string& fn)
{
    FILE* f = std::fopen(fn.c_str(), "r");
    if (!f)
        return;
    char key[160];
    int tcw, l1;
    double us;
    while (std::fscanf(f, "%159s %d %d %lf", key, &tcw, &l1, &us) == 4) {
        TuneChoice c;
        c.tile_cols = tcw;
        c.fft_l1 = l1;
        c.us = us;
        ctx->tune[key] = c;
    }
    std::fclose(f);
}

// Waits for the call and fills diag from the device triple (sample_grid_total,
// good count, gridding entries) at d_total.
void read_diag(stp_ctx* ctx, const stp_imager_pars* p, const double* d_total, stp_imager_diag* diag)
{
    unsigned long long h[3] = { 0, 0, 0 };
    HIP_CHECK(hipMemcpyAsync(h, d_total, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (!diag)
        return;
    double tot;
    std::memcpy(&tot, &h[0], sizeof(tot));
    diag->sample_grid_total = tot;
    diag->num_good_vis = h[1];
    diag->num_grid_entries = h[2];
    diag->image_rows = tot > 0.0 ? p->image_size : 0;
    diag->image_cols = tot > 0.0 ? p->image_size : 0;
    diag->ms_total = 0.0;
}

} // namespace

extern "C" {

stp_status stp_ctx_create(stp_ctx** out, int device)
{
    if (!out)
        return STP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return STP_ERR_DEVICE;
    if (device < 0 || device >= count)
        return STP_ERR_INVALID_ARGUMENT;
    stp_ctx* ctx = new (std::nothrow) stp_ctx();
    if (!ctx)
        return STP_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return STP_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        ctx->num_cus = prop.multiProcessorCount;
    if (const char* tf = std::getenv("STP_TUNE_FILE")) {
        ctx->tune_file = tf;
        load_tune_file(ctx, ctx->tune_file);
    }
    *out = ctx;
    return STP_OK;
}

stp_status stp_ctx_destroy(stp_ctx* ctx)
{
    if (!ctx)
        return STP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    invalidate_graphs(ctx);
    for (auto s : ctx->branch_streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (auto e : ctx->branch_events)
        (void)hipEventDestroy(e);
    if (ctx->copy_stream) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
    }
    if (ctx->copy_event)
        (void)hipEventDestroy(ctx->copy_event);
    if (ctx->skymodel_event)
        (void)hipEventDestroy(ctx->skymodel_event);
    for (auto& kv : ctx->batch_tab_events)
        (void)hipEventDestroy(kv.second);
    for (auto& kv : ctx->ws)
        if (kv.second.ptr)
            (void)hipFree(kv.second.ptr);
    for (auto& kv : ctx->host_ws)
        if (kv.second.ptr)
            (void)hipHostFree(kv.second.ptr);
    for (auto& kv : ctx->tables)
        if (kv.second.ptr)
            (void)hipFree(kv.second.ptr);
    for (auto& kv : ctx->twiddles) {
        (void)hipFree(kv.second.d);
        (void)hipFree(kv.second.f);
    }
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool)
        (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return STP_OK;
}

const char* stp_ctx_last_error(const stp_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

stp_status stp_ctx_synchronize(stp_ctx* ctx)
{
    return guarded(ctx, [&] { HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

stp_status stp_ctx_release_workspace(stp_ctx* ctx)
{
    return guarded(ctx, [&] {
        // only this context's queued work can still use its buffers
        sync_streams(ctx);
        invalidate_graphs(ctx);
        for (auto& kv : ctx->ws)
            if (kv.second.ptr)
                HIP_CHECK(hipFree(kv.second.ptr));
        ctx->ws.clear();
        for (auto& kv : ctx->tables)
            if (kv.second.ptr)
                HIP_CHECK(hipFree(kv.second.ptr));
        ctx->tables.clear();
        for (auto& kv : ctx->twiddles) {
            HIP_CHECK(hipFree(kv.second.d));
            HIP_CHECK(hipFree(kv.second.f));
        }
        ctx->twiddles.clear();
    });
}

size_t stp_ctx_workspace_bytes(const stp_ctx* ctx)
{
    if (!ctx)
        return 0;
    size_t s = 0;
    for (auto& kv : ctx->ws)
        s += kv.second.bytes;
    for (auto& kv : ctx->tables)
        s += kv.second.bytes;
    for (auto& kv : ctx->twiddles)
        s += size_t(kv.second.n) * (2 * sizeof(double) + 2 * sizeof(float));
    return s;
}

void stp_imager_pars_init(stp_imager_pars* p, uint32_t image_size, double cell_size)
{
    if (!p)
        return;
    // ImagerPars full constructor defaults (types.h:132-158)
    p->image_size = image_size;
    p->cell_size = cell_size;
    p->padding_factor = 1.0;
    p->padded_image_size = uint32_t(image_size * p->padding_factor);
    p->kernel_function = STP_KERNEL_PSWF;
    p->kernel_support = 3;
    p->kernel_exact = 1;
    p->oversampling = 1;
    p->generate_beam = 0;
    p->gridding_correction = 1;
    p->analytic_gcf = 1;
    p->r_fft = STP_FFT_ESTIMATE;
}

void stp_kernel_spec_init(stp_kernel_spec* k, int32_t kernel_function, double param)
{
    if (!k)
        return;
    k->kernel_function = kernel_function;
    k->param = param;
    // default width normalisations (conv_func.h:106-170, GaussianSinc :216-217)
    k->width_normalization_gaussian = (kernel_function == STP_KERNEL_GAUSSIANSINC) ? 2.52 : 1.0;
    k->width_normalization_sinc = (kernel_function == STP_KERNEL_GAUSSIANSINC) ? 1.55 : 1.0;
}

void stp_wproj_pars_init(stp_wproj_pars* p)
{
    if (!p)
        return;
    std::memset(p, 0, sizeof(*p));
    p->undersampling_opt = 1; // types.h:183-192
}

void stp_sourcefind_pars_init(stp_sourcefind_pars* p, double det, double ana)
{
    if (!p)
        return;
    // SourceFindImage constructor defaults (sourcefind.h:188-201)
    p->detection_n_sigma = det;
    p->analysis_n_sigma = ana;
    p->rms_est = 0.0;
    p->find_negative_sources = 1;
    p->sigma_clip_iters = 5;
    p->median_method = STP_MEDIAN_BINAPPROX;
    p->gaussian_fitting = 0;
    p->ccl_4connectivity = 0;
    p->generate_labelmap = 1;
    p->source_min_area = 5;
}

stp_status stp_image_visibilities_device(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp,
    const stp_aproj_pars* ap, const stp_kernel_spec* kernel, size_t n, const double* d_vis, const double* d_w, const double* d_uvw, int32_t precision,
    void* d_image, void* d_beam, int32_t synchronize, stp_imager_diag* diag)
{
    return guarded(ctx, [&] {
        check_imager_pars(p, wp, ap);
        STP_REQUIRE(d_image != nullptr, "image buffer is NULL");
        STP_REQUIRE(n == 0 || (d_vis && d_w && d_uvw), "visibility buffers are NULL");
        double* d_total = (double*)ws_get(ctx, "imager_total", sizeof(double) * 3);
        unsigned long long* d_ngood = (unsigned long long*)(d_total + 1);
        STP_REQUIRE(precision >= STP_PRECISION_FP64 && precision <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
        if (ctx->autotune)
            autotune_shape(ctx, p, wp && wp->num_wplanes > 0 ? wp : nullptr, kernel, n, d_vis, d_w, d_uvw, precision, d_image,
                d_beam, d_total, d_ngood);
        LaunchTimer whole(ctx, "image_visibilities", ctx->stream);
        run_imaging(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, precision, d_image, d_beam, d_total, d_ngood,
            (unsigned long long*)(d_total + 2));
        if (synchronize || diag)
            read_diag(ctx, p, d_total, diag);
    });
}

stp_status stp_image_visibilities_device_noalloc(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp,
    const stp_aproj_pars* ap, const stp_kernel_spec* kernel, size_t n, const double* d_vis, const double* d_w,
    const double* d_uvw, int32_t precision, void* d_image, void* d_beam, int32_t synchronize, stp_imager_diag* diag)
{
    return guarded(ctx, [&] {
        check_imager_pars(p, wp, ap);
        STP_REQUIRE(d_image != nullptr, "image buffer is NULL");
        STP_REQUIRE(n == 0 || (d_vis && d_w && d_uvw), "visibility buffers are NULL");
        STP_REQUIRE(precision >= STP_PRECISION_FP64 && precision <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
        char* base = (char*)scratch_region(ctx->device);
        // the previous call of this context may still read the region and the host
        // sources of its table uploads
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        Arena& ar = ctx->arena_state;
        ar.base = base;
        ar.cap = STP_NOALLOC_SCRATCH_BYTES;
        ar.used = 0;
        ar.slots.clear();
        ar.twiddles.clear();
        ar.keep_d.clear();
        ar.keep_f.clear();
        // launch timing would create events; it stays off for the call
        struct Scope {
            stp_ctx* c;
            bool timing;
            ~Scope()
            {
                c->arena = nullptr;
                c->timing = timing;
            }
        } scope{ ctx, ctx->timing };
        ctx->timing = false;
        ctx->arena = &ar;
        double* d_total = (double*)ws_get(ctx, "imager_total", sizeof(double) * 3);
        unsigned long long* d_ngood = (unsigned long long*)(d_total + 1);
        run_imaging(ctx, p, wp, kernel, n, d_vis, d_w, d_uvw, precision, d_image, d_beam, d_total, d_ngood,
            (unsigned long long*)(d_total + 2));
        if (synchronize || diag)
            read_diag(ctx, p, d_total, diag);
    });
}

stp_status stp_image_visibilities(stp_ctx* ctx, const stp_imager_pars* p, const stp_wproj_pars* wp,
    const stp_aproj_pars* ap, const stp_kernel_spec* kernel, size_t n, const double* vis, const double* w, const double* uvw, int32_t precision,
    double* image, double* beam, stp_imager_diag* diag)
{
    return guarded(ctx, [&] {
        check_imager_pars(p, wp, ap);
        STP_REQUIRE(image != nullptr, "image buffer is NULL");
        STP_REQUIRE(!p->generate_beam || beam != nullptr, "beam buffer is NULL with generate_beam");
        STP_REQUIRE(n == 0 || (vis && w && uvw), "visibility buffers are NULL");
        const size_t isz = size_t(p->image_size) * p->image_size;
        STP_REQUIRE(precision >= STP_PRECISION_FP64 && precision <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
        const size_t osz = (precision == STP_PRECISION_FP64) ? sizeof(double) : sizeof(float);
        double* d_in = (double*)ws_get(ctx, "host_in", sizeof(double) * (6 * n + 1));
        char* d_out = (char*)ws_get(ctx, "host_out", osz * isz * 2);
        if (n) {
            HIP_CHECK(hipMemcpyAsync(d_in, vis, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(d_in + 2 * n, w, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(d_in + 3 * n, uvw, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
        }
        stp_imager_diag dg{};
        const stp_status st = stp_image_visibilities_device(ctx, p, wp, ap, kernel, n, d_in, d_in + 2 * n, d_in + 3 * n,
            precision, d_out, p->generate_beam ? d_out + osz * isz : nullptr, 1, &dg);
        if (st != STP_OK)
            throw Error(st, ctx->last_error);
        if (dg.image_rows > 0) {
            if (precision == STP_PRECISION_FP64) {
                HIP_CHECK(hipMemcpyAsync(image, d_out, sizeof(double) * isz, hipMemcpyDeviceToHost, ctx->stream));
                if (p->generate_beam)
                    HIP_CHECK(hipMemcpyAsync(beam, d_out + osz * isz, sizeof(double) * isz, hipMemcpyDeviceToHost, ctx->stream));
                HIP_CHECK(hipStreamSynchronize(ctx->stream));
            } else {
                std::vector<float> tmp(isz);
                HIP_CHECK(hipStreamSynchronize(ctx->stream));
                HIP_CHECK(hipMemcpy(tmp.data(), d_out, sizeof(float) * isz, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < isz; i++)
                    image[i] = tmp[i];
                if (p->generate_beam) {
                    HIP_CHECK(hipMemcpy(tmp.data(), d_out + osz * isz, sizeof(float) * isz, hipMemcpyDeviceToHost));
                    for (size_t i = 0; i < isz; i++)
                        beam[i] = tmp[i];
                }
            }
        }
        if (diag)
            *diag = dg;
    });
}

stp_status stp_convolve_to_grid(stp_ctx* ctx, const stp_kernel_spec* kernel, uint32_t support,
    int32_t image_size, size_t n, const double* uv, const double* vis, const double* w, int32_t kernel_exact,
    uint32_t oversampling, int32_t shift_uv, int32_t halfplane, int32_t generate_beam, double* vis_grid,
    double* sampling_grid, double* sample_grid_total)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(vis_grid != nullptr, "vis_grid is NULL");
        STP_REQUIRE(!generate_beam || sampling_grid != nullptr, "sampling_grid is NULL with generate_beam");
        STP_REQUIRE(kernel != nullptr, "kernel spec is NULL");
        check_kernel_spec(kernel);
        STP_REQUIRE(kernel_exact || oversampling >= 1, "oversampling must be >= 1");
        STP_REQUIRE(n == 0 || (uv && vis && w), "visibility buffers are NULL");
        const int N = image_size;
        const int rows = halfplane ? N / 2 + 1 : N;
        const size_t ge = size_t(rows) * size_t(N);
        double* d_in = (double*)ws_get(ctx, "host_in", sizeof(double) * (5 * n + 1));
        if (n) {
            HIP_CHECK(hipMemcpyAsync(d_in, uv, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(d_in + 2 * n, vis, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
            HIP_CHECK(hipMemcpyAsync(d_in + 4 * n, w, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        }
        cplx<double>* grids = (cplx<double>*)ws_get(ctx, "unit_grids", sizeof(cplx<double>) * ge * 2);
        double* d_total = (double*)ws_get(ctx, "unit_total", sizeof(double));
        GridderArgs a;
        a.image_size = N;
        a.support = int(support);
        a.oversampling = int(oversampling);
        a.kernel_exact = kernel_exact != 0;
        a.shift_uv = shift_uv != 0;
        a.halfplane = halfplane != 0;
        a.generate_beam = generate_beam != 0;
        a.uv_scale = 1.0;
        a.kernel = *kernel;
        grid_visibilities<double>(ctx, a, n, d_in, d_in + 2 * n, d_in + 4 * n, grids, generate_beam ? grids + ge : nullptr,
            d_total, nullptr);
        HIP_CHECK(hipMemcpyAsync(vis_grid, grids, sizeof(cplx<double>) * ge, hipMemcpyDeviceToHost, ctx->stream));
        if (generate_beam)
            HIP_CHECK(hipMemcpyAsync(sampling_grid, grids + ge, sizeof(cplx<double>) * ge, hipMemcpyDeviceToHost, ctx->stream));
        double tot = 0.0;
        HIP_CHECK(hipMemcpyAsync(&tot, d_total, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (sample_grid_total)
            *sample_grid_total = tot;
    });
}

stp_status stp_fft_c2r(stp_ctx* ctx, const double* in, int32_t in_rows, int32_t n_cols, double* out)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(in && out, "NULL buffer");
        STP_REQUIRE(in_rows >= 3, "in_rows must be n/2+1 with n >= 4");
        const int n = 2 * (in_rows - 1);
        const size_t ge = size_t(in_rows) * n_cols;
        cplx<double>* g = (cplx<double>*)ws_get(ctx, "unit_fft_in", sizeof(cplx<double>) * ge);
        double* o = (double*)ws_get(ctx, "unit_fft_out", sizeof(double) * size_t(n) * n_cols);
        HIP_CHECK(hipMemcpyAsync(g, in, sizeof(cplx<double>) * ge, hipMemcpyHostToDevice, ctx->stream));
        // the output is n x n_cols; the last pass writes columns through the imager's
        // crop mapping with image_size = n, the identity for n_cols <= n
        STP_REQUIRE(n_cols <= n, "c2r needs n_cols <= n (rows of the real output)");
        NormSpec ns;
        ns.image_size = n;
        OutPtrs<double> outs{};
        outs.p[0] = o;
        fft_c2r_2d<double, double>(ctx, g, 1, n, n_cols, outs, ns);
        HIP_CHECK(hipMemcpyAsync(out, o, sizeof(double) * size_t(n) * n_cols, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_fft_c2c(stp_ctx* ctx, const double* in, int32_t n, int32_t forward, double* out)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(in && out, "NULL buffer");
        const size_t ne = size_t(n) * n;
        cplx<double>* a = (cplx<double>*)ws_get(ctx, "unit_c2c_a", sizeof(cplx<double>) * ne);
        cplx<double>* b = (cplx<double>*)ws_get(ctx, "unit_c2c_b", sizeof(cplx<double>) * ne);
        HIP_CHECK(hipMemcpyAsync(a, in, sizeof(cplx<double>) * ne, hipMemcpyHostToDevice, ctx->stream));
        fft_c2c_2d<double>(ctx, a, b, n, forward != 0);
        HIP_CHECK(hipMemcpyAsync(out, b, sizeof(cplx<double>) * ne, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_generate_visibilities_from_local_skymodel(stp_ctx* ctx, size_t n_src, const double* skymodel, size_t n_vis,
    const double* uvw, double* model_vis)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(n_vis == 0 || (uvw && model_vis), "visibility buffers are NULL");
        if (n_vis == 0)
            return;
        double* d = (double*)ws_get(ctx, "host_model", sizeof(double) * 5 * n_vis);
        HIP_CHECK(hipMemcpyAsync(d + 2 * n_vis, uvw, sizeof(double) * 3 * n_vis, hipMemcpyHostToDevice, ctx->stream));
        model_visibilities(ctx, n_src, skymodel, n_vis, d + 2 * n_vis, d, false);
        HIP_CHECK(hipMemcpyAsync(model_vis, d, sizeof(double) * 2 * n_vis, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_model_visibilities_device(stp_ctx* ctx, size_t n_src, const double* skymodel, size_t n_vis,
    const double* d_uvw, double* d_vis, int32_t subtract, int32_t synchronize)
{
    return guarded(ctx, [&] {
        model_visibilities(ctx, n_src, skymodel, n_vis, d_uvw, d_vis, subtract != 0);
        if (synchronize)
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_source_find(stp_ctx* ctx, const double* image, int32_t rows, int32_t cols, const stp_sourcefind_pars* p,
    int32_t* label_map, stp_island* islands, size_t cap, size_t* n_islands, stp_sourcefind_stats* stats)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(image && p && n_islands, "NULL argument");
        STP_REQUIRE(rows > 0 && cols > 0, "empty image");
        const size_t ne = size_t(rows) * cols;
        double* d_img = (double*)ws_get(ctx, "host_sf_img", sizeof(double) * ne);
        int32_t* d_lab = label_map ? (int32_t*)ws_get(ctx, "host_sf_lab", sizeof(int32_t) * ne) : nullptr;
        HIP_CHECK(hipMemcpyAsync(d_img, image, sizeof(double) * ne, hipMemcpyHostToDevice, ctx->stream));
        source_find_device(ctx, d_img, STP_PRECISION_FP64, rows, cols, *p, d_lab, islands, cap, n_islands, stats);
        if (label_map) {
            HIP_CHECK(hipMemcpyAsync(label_map, d_lab, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, ctx->stream));
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

stp_status stp_source_find_device(stp_ctx* ctx, const void* d_image, int32_t prec, int32_t rows, int32_t cols,
    const stp_sourcefind_pars* p, int32_t* d_label_map, stp_island* islands, size_t cap, size_t* n_islands,
    stp_sourcefind_stats* stats)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(d_image && p && n_islands, "NULL argument");
        STP_REQUIRE(prec >= STP_PRECISION_FP64 && prec <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
        STP_REQUIRE(cap == 0 || islands != nullptr, "island buffer is NULL with a non-zero capacity");
        source_find_device(ctx, d_image, prec, rows, cols, *p, d_label_map, islands, cap, n_islands, stats);
    });
}

stp_status stp_fit_islands_device(stp_ctx* ctx, const void* d_image, int32_t prec, int32_t rows, int32_t cols,
    const int32_t* d_label_map, const stp_island* islands, size_t n, stp_fit_result* results)
{
    return guarded(ctx, [&] { fit_islands_device(ctx, d_image, prec, rows, cols, d_label_map, islands, n, results); });
}

stp_status stp_fit_islands(stp_ctx* ctx, const double* image, int32_t rows, int32_t cols, const int32_t* label_map,
    const stp_island* islands, size_t n, stp_fit_result* results)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(rows >= 2 && cols >= 2, "image must be at least 2 x 2");
        STP_REQUIRE(image && label_map, "image or label map is NULL");
        if (n == 0)
            return;
        const size_t ne = size_t(rows) * cols;
        double* d_img = (double*)ws_get(ctx, "host_fit_img", sizeof(double) * ne);
        int32_t* d_lab = (int32_t*)ws_get(ctx, "host_fit_lab", sizeof(int32_t) * ne);
        HIP_CHECK(hipMemcpyAsync(d_img, image, sizeof(double) * ne, hipMemcpyHostToDevice, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(d_lab, label_map, sizeof(int32_t) * ne, hipMemcpyHostToDevice, ctx->stream));
        fit_islands_device(ctx, d_img, STP_PRECISION_FP64, rows, cols, d_lab, islands, n, results);
    });
}

stp_status stp_device_malloc(stp_ctx* ctx, void** ptr, size_t bytes)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(ptr != nullptr, "NULL pointer");
        HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
    });
}

stp_status stp_host_malloc(stp_ctx* ctx, void** ptr, size_t bytes)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(ptr != nullptr, "NULL pointer");
        HIP_CHECK(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    });
}

stp_status stp_host_free(stp_ctx* ctx, void* ptr)
{
    return guarded(ctx, [&] {
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->copy_stream)
            HIP_CHECK(hipStreamSynchronize(ctx->copy_stream));
        HIP_CHECK(hipHostFree(ptr));
    });
}

stp_status stp_device_free(stp_ctx* ctx, void* ptr)
{
    return guarded(ctx, [&] {
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipFree(ptr));
    });
}

stp_status stp_memcpy_htod(stp_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return guarded(ctx, [&] {
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_memcpy_dtoh(stp_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return guarded(ctx, [&] {
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

stp_status stp_ctx_set_timing(stp_ctx* ctx, int32_t enable)
{
    return guarded(ctx, [&] { ctx->timing = enable != 0; });
}

stp_status stp_ctx_set_autotune(stp_ctx* ctx, int32_t mode, const char* tune_file)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(mode == 0 || mode == 1, "autotune mode must be 0 or 1");
        ctx->autotune = mode;
        ctx->tune_file = tune_file ? tune_file : "";
        if (!ctx->tune_file.empty())
            load_tune_file(ctx, ctx->tune_file);
    });
}

stp_status stp_ctx_get_tuning(stp_ctx* ctx, const stp_imager_pars* img_pars, const stp_wproj_pars* w_proj, size_t n_vis,
    int32_t precision, int32_t* tile_cols, int32_t* fft_l1, double* us)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(img_pars != nullptr, "img_pars is NULL");
        auto it = ctx->tune.find(tune_key(img_pars, w_proj && w_proj->num_wplanes > 0 ? w_proj : nullptr, n_vis, precision));
        const TuneChoice c = it != ctx->tune.end() ? it->second : TuneChoice{};
        if (tile_cols)
            *tile_cols = c.tile_cols;
        if (fft_l1)
            *fft_l1 = c.fft_l1;
        if (us)
            *us = c.us;
    });
}

stp_status stp_image_batch(stp_ctx* ctx, size_t n_items, const stp_batch_item* items, const stp_kernel_spec* kernel,
    int32_t precision, uint32_t max_concurrent, uint32_t flags, double* sample_grid_totals)
{
    return guarded(ctx, [&] {
        STP_REQUIRE(n_items == 0 || items != nullptr, "items is NULL");
        STP_REQUIRE(precision >= STP_PRECISION_FP64 && precision <= STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
        for (size_t i = 0; i < n_items; i++) {
            const stp_batch_item& it = items[i];
            check_imager_pars(&it.img_pars, &it.w_proj, nullptr);
            STP_REQUIRE(it.d_image != nullptr, "batch item image buffer is NULL");
            STP_REQUIRE(!it.img_pars.generate_beam || it.d_beam != nullptr, "batch item beam buffer is NULL");
            STP_REQUIRE(it.n_vis == 0 || (it.d_vis && it.d_vis_weights && it.d_uvw_lambda), "batch item visibilities are NULL");
        }
        if (n_items == 0)
            return;
        const stp_wproj_pars* none = nullptr;
        auto wp_of = [&](const stp_batch_item& it) { return it.w_proj.num_wplanes > 0 ? &it.w_proj : none; };
        // per-snapshot totals and good counts live in the base workspace
        double* totals = (double*)ws_get(ctx, "batch_totals", sizeof(double) * 2 * n_items);
        if ((flags & STP_BATCH_AUTOTUNE) || ctx->autotune)
            for (size_t i = 0; i < n_items; i++) {
                const stp_batch_item& it = items[i];
                autotune_shape(ctx, &it.img_pars, wp_of(it), kernel, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                    precision, it.d_image, it.d_beam, totals + 2 * i, (unsigned long long*)(totals + 2 * i + 1));
            }
        // fused snapshots (one launch per stage, main stream) and the rest (streams)
        std::vector<size_t> fused, rest;
        for (size_t i = 0; i < n_items; i++)
            ((flags & STP_BATCH_FUSED) && batch_fusable(ctx, items[i], kernel, precision) ? fused : rest).push_back(i);
        hipStream_t main = ctx->stream;
        // streams: K for the rest, two more for the fused snapshots' chunks
        const int K = rest.empty() ? 0
                                   : int(std::min<size_t>(rest.size(), max_concurrent ? std::min(max_concurrent, 16u) : 4u));
        while (int(ctx->branch_streams.size()) < K + 2) {
            hipStream_t st;
            HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            ctx->branch_streams.push_back(st);
        }
        while (ctx->branch_events.size() < size_t(K) + 7) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ctx->branch_events.push_back(e);
        }
        auto run_fused = [&]() {
            if (fused.empty())
                return;
            // chunks of about equal work (visibilities, and grid cells for the FFT)
            const int P = fused.size() >= 8 ? 4 : (fused.size() >= 4 ? 2 : 1);
            std::vector<std::vector<size_t>> part(P);
            auto cost = [&](size_t i) {
                const double N = double(items[i].img_pars.padded_image_size);
                return double(items[i].n_vis) + 1.1e6 * (N / 4096.0) * (N / 4096.0);
            };
            std::vector<size_t> ord(fused);
            std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
            std::vector<double> load(P, 0.0);
            for (size_t i : ord) {
                const int j = int(std::min_element(load.begin(), load.end()) - load.begin());
                part[j].push_back(i);
                load[j] += cost(i);
            }
            for (auto& pt : part)
                std::sort(pt.begin(), pt.end());
            const hipStream_t* st = &ctx->branch_streams[K];
            const hipEvent_t* ev = &ctx->branch_events[K + 3];
            HIP_CHECK(hipEventRecord(ctx->branch_events[K + 1], main));
            for (int s = 0; s < 2; s++)
                HIP_CHECK(hipStreamWaitEvent(st[s], ctx->branch_events[K + 1], 0));
            try {
                if (precision == STP_PRECISION_FP64)
                    run_fused_chunks<double, double>(ctx, items, part, kernel, precision, totals, st, ev, K + 1);
                else
                    run_fused_chunks<float, float>(ctx, items, part, kernel, precision, totals, st, ev, K + 1);
            } catch (...) {
                ctx->stream = main;
                ctx->ws_slot = 0;
                throw;
            }
            ctx->stream = main;
            ctx->ws_slot = 0;
            for (int s = 0; s < 2; s++) {
                HIP_CHECK(hipEventRecord(ctx->branch_events[K + 2], st[s]));
                HIP_CHECK(hipStreamWaitEvent(main, ctx->branch_events[K + 2], 0));
            }
        };
        auto finish = [&]() {
            if (sample_grid_totals) {
                std::vector<double> t(2 * n_items);
                HIP_CHECK(hipMemcpyAsync(t.data(), totals, sizeof(double) * 2 * n_items, hipMemcpyDeviceToHost, main));
                HIP_CHECK(hipStreamSynchronize(main));
                for (size_t i = 0; i < n_items; i++)
                    sample_grid_totals[i] = t[2 * i];
            } else {
                HIP_CHECK(hipStreamSynchronize(main));
            }
        };
        if (rest.empty()) {
            run_fused();
            finish();
            return;
        }
        auto enqueue = [&](bool with_fused) {
            HIP_CHECK(hipEventRecord(ctx->branch_events[K], main));
            for (int b = 0; b < K; b++)
                HIP_CHECK(hipStreamWaitEvent(ctx->branch_streams[b], ctx->branch_events[K], 0));
            try {
                for (size_t r = 0; r < rest.size(); r++) {
                    const int b = int(r % size_t(K));
                    ctx->stream = ctx->branch_streams[b];
                    ctx->ws_slot = b + 1;
                    const size_t i = rest[r];
                    const stp_batch_item& it = items[i];
                    run_imaging(ctx, &it.img_pars, wp_of(it), kernel, it.n_vis, it.d_vis, it.d_vis_weights, it.d_uvw_lambda,
                        precision, it.d_image, it.d_beam, totals + 2 * i, (unsigned long long*)(totals + 2 * i + 1));
                }
            } catch (...) {
                ctx->stream = main;
                ctx->ws_slot = 0;
                throw;
            }
            ctx->stream = main;
            ctx->ws_slot = 0;
            if (with_fused)
                run_fused(); // on the main stream, beside the branches
            for (int b = 0; b < K; b++) {
                HIP_CHECK(hipEventRecord(ctx->branch_events[b], ctx->branch_streams[b]));
                HIP_CHECK(hipStreamWaitEvent(main, ctx->branch_events[b], 0));
            }
        };
        if (flags & STP_BATCH_GRAPH) {
            // signature: everything the captured launches depend on
            std::string sig = "K" + std::to_string(K) + ":p" + std::to_string(precision);
            char buf[512];
            if (kernel) {
                snprintf(buf, sizeof(buf), ":k%d,%.17g,%.17g,%.17g", kernel->kernel_function, kernel->param,
                    kernel->width_normalization_gaussian, kernel->width_normalization_sinc);
                sig += buf;
            }
            for (size_t i : rest) {
                const stp_batch_item& it = items[i];
                const stp_imager_pars& q = it.img_pars;
                snprintf(buf, sizeof(buf), "|%u,%u,%.17g,%.17g,%d,%u,%d,%u,%d,%d,%d,%u,%u,%u,%.17g,%d,%zu,%p,%p,%p,%p,%p", q.image_size,
                    q.padded_image_size, q.cell_size, q.padding_factor, q.kernel_function, q.kernel_support, q.kernel_exact,
                    q.oversampling, q.generate_beam, q.gridding_correction, q.analytic_gcf, it.w_proj.num_wplanes,
                    it.w_proj.max_wpconv_support, it.w_proj.undersampling_opt, it.w_proj.kernel_trunc_perc,
                    it.w_proj.wplanes_median, it.n_vis, (const void*)it.d_vis, (const void*)it.d_vis_weights,
                    (const void*)it.d_uvw_lambda, it.d_image, it.d_beam);
                sig += buf;
            }
            // the fused snapshots first, eagerly (their tables change with every batch)
            run_fused();
            auto g = ctx->graphs.find(sig);
            if (g != ctx->graphs.end()) {
                HIP_CHECK(hipGraphLaunch(g->second, main));
            } else {
                // eager run: sizes every workspace and uploads every table, so the
                // capture below records kernel launches only
                enqueue(false);
                HIP_CHECK(hipStreamSynchronize(main));
                const bool timing = ctx->timing;
                ctx->timing = false;
                hipGraph_t graph = nullptr;
                HIP_CHECK(hipStreamBeginCapture(main, hipStreamCaptureModeRelaxed));
                try {
                    enqueue(false);
                } catch (...) {
                    hipGraph_t dropped = nullptr;
                    (void)hipStreamEndCapture(main, &dropped);
                    if (dropped)
                        (void)hipGraphDestroy(dropped);
                    ctx->timing = timing;
                    throw;
                }
                ctx->timing = timing;
                HIP_CHECK(hipStreamEndCapture(main, &graph));
                hipGraphExec_t exec = nullptr;
                const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_CHECK(ie);
                // a bounded cache: batches with new buffers or sizes on every call would
                // otherwise keep one graph each; the oldest goes once its launches are done
                if (ctx->graphs.size() >= MAX_BATCH_GRAPHS) {
                    HIP_CHECK(hipStreamSynchronize(main));
                    auto old = ctx->graphs.find(ctx->graph_order.front());
                    if (old != ctx->graphs.end()) {
                        (void)hipGraphExecDestroy(old->second);
                        ctx->graphs.erase(old);
                    }
                    ctx->graph_order.pop_front();
                }
                ctx->graphs[sig] = exec;
                ctx->graph_order.push_back(sig);
            }
        } else {
            enqueue(true);
        }
        finish();
    });
}

stp_status stp_ctx_set_timing_filter(stp_ctx* ctx, const char* names)
{
    return guarded(ctx, [&] {
        ctx->timing_filter.clear();
        if (!names)
            return;
        std::string cur;
        for (const char* c = names;; c++) {
            if (*c == ',' || *c == 0) {
                if (!cur.empty())
                    ctx->timing_filter.push_back(cur);
                cur.clear();
                if (*c == 0)
                    break;
            } else {
                cur.push_back(*c);
            }
        }
    });
}

stp_status stp_ctx_reset_timings(stp_ctx* ctx)
{
    return guarded(ctx, [&] {
        resolve_timings(ctx);
        ctx->timings.clear();
    });
}

stp_status stp_ctx_get_timings(stp_ctx* ctx, char* names, double* total_ms, uint64_t* counts, size_t cap, size_t* n)
{
    return guarded(ctx, [&] {
        resolve_timings(ctx);
        size_t i = 0;
        for (auto& kv : ctx->timings) {
            if (i < cap) {
                if (names) {
                    std::memset(names + 64 * i, 0, 64);
                    std::strncpy(names + 64 * i, kv.first.c_str(), 63);
                }
                if (total_ms)
                    total_ms[i] = kv.second.ms;
                if (counts)
                    counts[i] = kv.second.count;
            }
            i++;
        }
        if (n)
            *n = i;
    });
}

} // extern "C"
