// Context, workspace, twiddle tables and launch timing of the gfx950 STP library.
#include <cmath>
#include <cstring>

#include "common.h"

namespace stp_hip {

// Next 256-byte aligned piece of the no-allocation region.
static void* arena_carve(Arena* ar, size_t bytes)
{
    const size_t off = (ar->used + 255) & ~size_t(255);
    if (off > ar->cap || bytes > ar->cap - off)
        throw Error(STP_ERR_CAPACITY, "no-allocation call needs more than the " + std::to_string(ar->cap) +
                "-byte device scratch region (" + std::to_string(off + bytes) + " bytes so far)");
    ar->used = off + bytes;
    return ar->base + off;
}

void* ws_get(stp_ctx* ctx, const std::string& name, size_t bytes)
{
    if (bytes == 0)
        bytes = 256;
    if (ctx->arena) {
        DevBuf& s = ctx->arena->slots[name];
        if (s.bytes < bytes) {
            s.ptr = arena_carve(ctx->arena, bytes);
            s.bytes = bytes;
        }
        return s.ptr;
    }
    DevBuf& b = ctx->ws[ctx->ws_slot ? name + "#" + std::to_string(ctx->ws_slot) : name];
    if (b.bytes >= bytes)
        return b.ptr;
    if (b.ptr) {
        // the old buffer may still be read by this context's queued work, and
        // captured batch graphs refer to it
        sync_streams(ctx);
        HIP_CHECK(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
        invalidate_graphs(ctx);
    }
    const size_t alloc = (bytes + 4095) & ~size_t(4095);
    HIP_CHECK(hipMalloc(&b.ptr, alloc));
    b.bytes = alloc;
    return b.ptr;
}

void sync_streams(stp_ctx* ctx)
{
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto st : ctx->branch_streams)
        HIP_CHECK(hipStreamSynchronize(st));
    if (ctx->copy_stream)
        HIP_CHECK(hipStreamSynchronize(ctx->copy_stream));
}

hipStream_t copy_stream(stp_ctx* ctx)
{
    if (!ctx->copy_stream) {
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->copy_event, hipEventDisableTiming));
    }
    return ctx->copy_stream;
}

void* pinned_device_ptr(void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // ordinary (pageable) memory
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost)
        return nullptr;
    void* d = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(&d, p, 0));
    return d;
}

size_t ws_bytes(stp_ctx* ctx, const std::string& name)
{
    auto it = ctx->ws.find(ctx->ws_slot ? name + "#" + std::to_string(ctx->ws_slot) : name);
    return it == ctx->ws.end() ? 0 : it->second.bytes;
}

void invalidate_graphs(stp_ctx* ctx)
{
    for (auto& kv : ctx->graphs)
        (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
    ctx->graph_order.clear();
}

void* host_stage(stp_ctx* ctx, const std::string& name, size_t bytes)
{
    if (bytes == 0)
        bytes = 256;
    if (ctx->arena)
        throw Error(STP_ERR_INTERNAL, "host staging requested inside a no-allocation call");
    DevBuf& b = ctx->host_ws[name];
    if (b.bytes >= bytes)
        return b.ptr;
    if (b.ptr) {
        // queued copies may still target the old buffer
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipHostFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    const size_t alloc = (bytes + 4095) & ~size_t(4095);
    HIP_CHECK(hipHostMalloc(&b.ptr, alloc, hipHostMallocDefault));
    b.bytes = alloc;
    return b.ptr;
}

bool has_table(stp_ctx* ctx, const std::string& key)
{
    return ctx->tables.count(key) != 0 || (ctx->arena && ctx->arena->slots.count("table:" + key) != 0);
}

const double* dev_table(stp_ctx* ctx, const std::string& key, const std::vector<double>& host)
{
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end())
        return (const double*)it->second.ptr;
    if (ctx->arena) {
        // tables the context does not hold yet are uploaded into the region, on the
        // context stream, for this call only
        auto at = ctx->arena->slots.find("table:" + key);
        if (at != ctx->arena->slots.end())
            return (const double*)at->second.ptr;
        DevBuf s;
        s.bytes = std::max<size_t>(8, host.size() * sizeof(double));
        s.ptr = arena_carve(ctx->arena, s.bytes);
        if (!host.empty()) {
            ctx->arena->keep_d.push_back(host);
            HIP_CHECK(hipMemcpyAsync(s.ptr, ctx->arena->keep_d.back().data(), host.size() * sizeof(double),
                hipMemcpyHostToDevice, ctx->stream));
        }
        ctx->arena->slots["table:" + key] = s;
        return (const double*)s.ptr;
    }
    DevBuf b;
    b.bytes = std::max<size_t>(8, host.size() * sizeof(double));
    HIP_CHECK(hipMalloc(&b.ptr, b.bytes));
    if (!host.empty())
        HIP_CHECK(hipMemcpy(b.ptr, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->tables[key] = b;
    return (const double*)b.ptr;
}

// exp(+2 pi i k / n) for k < n, evaluated in long double and rounded once.
static void twiddle_tables(int n, std::vector<double>& hd, std::vector<float>& hf)
{
    hd.resize(2 * size_t(n));
    hf.resize(2 * size_t(n));
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; k++) {
        const long double a = two_pi * (long double)k / (long double)n;
        const long double c = cosl(a), s = sinl(a);
        hd[2 * k] = double(c);
        hd[2 * k + 1] = double(s);
        hf[2 * k] = float(c);
        hf[2 * k + 1] = float(s);
    }
}

const Twiddles& twiddles_for(stp_ctx* ctx, int n)
{
    auto it = ctx->twiddles.find(n);
    if (it != ctx->twiddles.end())
        return it->second;
    std::vector<double> hd;
    std::vector<float> hf;
    if (ctx->arena) {
        Arena* ar = ctx->arena;
        auto at = ar->twiddles.find(n);
        if (at != ar->twiddles.end())
            return at->second;
        twiddle_tables(n, hd, hf);
        Twiddles t;
        t.n = n;
        t.d = (double*)arena_carve(ar, hd.size() * sizeof(double));
        t.f = (float*)arena_carve(ar, hf.size() * sizeof(float));
        ar->keep_d.push_back(std::move(hd));
        ar->keep_f.push_back(std::move(hf));
        HIP_CHECK(hipMemcpyAsync(t.d, ar->keep_d.back().data(), ar->keep_d.back().size() * sizeof(double),
            hipMemcpyHostToDevice, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(t.f, ar->keep_f.back().data(), ar->keep_f.back().size() * sizeof(float),
            hipMemcpyHostToDevice, ctx->stream));
        return ar->twiddles.emplace(n, t).first->second;
    }
    twiddle_tables(n, hd, hf);
    Twiddles t;
    t.n = n;
    HIP_CHECK(hipMalloc(&t.d, hd.size() * sizeof(double)));
    HIP_CHECK(hipMalloc(&t.f, hf.size() * sizeof(float)));
    HIP_CHECK(hipMemcpyAsync(t.d, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(t.f, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return ctx->twiddles.emplace(n, t).first->second;
}

static hipEvent_t take_event(stp_ctx* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

LaunchTimer::LaunchTimer(stp_ctx* c, const char* n, hipStream_t s)
    : ctx(c)
    , name(n)
    , stream(s)
{
    bool on = ctx->timing;
    if (on && !ctx->timing_filter.empty()) {
        on = false;
        for (const auto& f : ctx->timing_filter)
            if (f == n) {
                on = true;
                break;
            }
    }
    if (on) {
        a = take_event(ctx);
        b = take_event(ctx);
        HIP_CHECK(hipEventRecord(a, stream));
    }
}

LaunchTimer::~LaunchTimer()
{
    if (a) {
        if (hipEventRecord(b, stream) == hipSuccess)
            ctx->pending.push_back({ name, a, b });
    }
}

void resolve_timings(stp_ctx* ctx)
{
    if (ctx->pending.empty())
        return;
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        TimingAcc& acc = ctx->timings[p.name];
        acc.ms += ms;
        acc.count++;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
}

} // namespace stp_hip
