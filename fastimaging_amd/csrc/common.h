// Shared definitions of the gfx950 STP library: context, workspace, errors,
// launch timing and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <deque>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/stp_capi.h"

namespace stp_hip {

// Error carried from the implementation to the C-ABI boundary.
struct Error : std::runtime_error {
    stp_status code;
    Error(stp_status c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

#define STP_REQUIRE(cond, msg)                                              \
    do {                                                                    \
        if (!(cond))                                                        \
            throw ::stp_hip::Error(STP_ERR_INVALID_ARGUMENT, (msg));        \
    } while (0)

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw ::stp_hip::Error(e_ == hipErrorOutOfMemory ? STP_ERR_OUT_OF_MEMORY : STP_ERR_DEVICE, \
                std::string(#expr) + ": " + hipGetErrorString(e_));                              \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct TimingAcc {
    double ms = 0.0;
    uint64_t count = 0;
};

// Launch configuration picked by the autotuner for one imaging shape.
struct TuneChoice {
    int tile_cols = 0; // gridder tile width
    int fft_l1 = 0;    // four-step split of the u-axis FFT
    double us = 0.0;   // measured imaging time with this choice
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};

// Twiddle table exp(+2 pi i k / n), k < n, in double and float.
struct Twiddles {
    double* d = nullptr; // 2n doubles (re, im)
    float* f = nullptr;  // 2n floats
    int n = 0;
};

// Fixed device region that a no-allocation call carves its workspace, kernel
// tables and twiddles from (bump allocation, 256-byte aligned), in place of the
// context's grow-only buffers.  Set for the duration of one call.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    std::map<std::string, DevBuf> slots;       // workspace and table slots of this call
    std::map<int, Twiddles> twiddles;          // twiddle tables uploaded by this call
    std::vector<std::vector<double>> keep_d;   // host sources of queued table uploads
    std::vector<std::vector<float>> keep_f;
};

} // namespace stp_hip

struct stp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::map<std::string, stp_hip::DevBuf> ws;
    std::map<std::string, stp_hip::DevBuf> host_ws; // pinned host staging (hipHostMalloc)
    std::map<int, stp_hip::Twiddles> twiddles;
    std::map<std::string, stp_hip::DevBuf> tables;
    // non-null while a no-allocation call runs: every buffer comes from it.  The
    // state outlives the call so that host sources of queued uploads stay valid
    // until the next no-allocation call has waited for the stream.
    stp_hip::Arena* arena = nullptr;
    stp_hip::Arena arena_state;
    bool timing = false;
    std::vector<std::string> timing_filter; // empty: every launch
    // autotuner: chosen launch configuration per imaging shape; mode 0 off,
    // 1 tune shapes on first use; results appended to tune_file when set
    std::map<std::string, stp_hip::TuneChoice> tune;
    int autotune = 0;
    std::string tune_file;
    // batches: workspace namespace of the running branch and the branch streams
    int ws_slot = 0;
    std::vector<hipStream_t> branch_streams;
    std::vector<hipEvent_t> branch_events;
    std::map<std::string, hipGraphExec_t> graphs;
    std::deque<std::string> graph_order; // capture order: the oldest graph is dropped beyond MAX_BATCH_GRAPHS
    std::vector<stp_hip::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, stp_hip::TimingAcc> timings;
    int num_cus = 256;
    // islands found by the last source find: the size of the island copy enqueued
    // with the final state copy (one host round trip when it suffices)
    size_t sf_islands_hint = 64;
    // second stream for result copies that overlap the last kernels of a call
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_event = nullptr;
    // the last upload of sky-model terms from the page-locked stage
    hipEvent_t skymodel_event = nullptr;
    // per workspace slot: the last upload of a fused batch's tables from its page-locked stage
    std::map<int, hipEvent_t> batch_tab_events;
};

namespace stp_hip {

// Returns a device buffer of at least `bytes` for slot `name` (contents undefined).
void* ws_get(stp_ctx* ctx, const std::string& name, size_t bytes);
// Waits for every stream of the context (main, batch branches, copy stream).
void sync_streams(stp_ctx* ctx);
// The context's copy stream (created on first use, with ctx->copy_event).
hipStream_t copy_stream(stp_ctx* ctx);
// The device address of host pointer p when it lies in page-locked memory of the
// HIP runtime (kernels may write it directly), else null.
void* pinned_device_ptr(void* p);
// Bytes held by workspace slot `name` (0 when it has none).
size_t ws_bytes(stp_ctx* ctx, const std::string& name);
// Destroys the context's captured batch graphs (their buffers are about to change).
void invalidate_graphs(stp_ctx* ctx);
// Returns a pinned host buffer of at least `bytes` for slot `name` (contents undefined).
void* host_stage(stp_ctx* ctx, const std::string& name, size_t bytes);
// Device copy of a host table, uploaded once per key and kept for the context's lifetime.
const double* dev_table(stp_ctx* ctx, const std::string& key, const std::vector<double>& host);
bool has_table(stp_ctx* ctx, const std::string& key);
const Twiddles& twiddles_for(stp_ctx* ctx, int n);

// Kernel-launch bracket: records HIP events on the context stream when timing is on.
struct LaunchTimer {
    stp_ctx* ctx;
    const char* name;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchTimer(stp_ctx* c, const char* n, hipStream_t s);
    ~LaunchTimer();
};

void resolve_timings(stp_ctx* ctx);

// Device address of the no-allocation scratch region (STP_NOALLOC_SCRATCH_BYTES)
// on `device`, the current device; resolved once per device.
void* scratch_region(int device);

template <typename K, typename... Args>
inline void launch(stp_ctx* ctx, const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args)
{
    if (grid.x == 0 || grid.y == 0 || grid.z == 0)
        return;
    LaunchTimer t(ctx, name, ctx->stream);
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
    HIP_CHECK(hipGetLastError());
}

// The same on another stream of the context (its copy stream).
template <typename K, typename... Args>
inline void launch_on(stp_ctx* ctx, hipStream_t stream, const char* name, K kernel, dim3 grid, dim3 block, size_t shmem,
    Args... args)
{
    if (grid.x == 0 || grid.y == 0 || grid.z == 0)
        return;
    LaunchTimer t(ctx, name, stream);
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
    HIP_CHECK(hipGetLastError());
}

inline unsigned div_up(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

} // namespace stp_hip

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return { a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x };
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return { a.x + b.x, a.y + b.y }; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return { a.x - b.x, a.y - b.y }; }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return { a.x, -a.y }; }
