// Canonical-order pair reduction and exclusive scan kernels (see reduce.h).
#include <algorithm>

#include "reduce.h"

#pragma clang fp contract(off)

namespace {

using namespace stp_hip;

__global__ __launch_bounds__(CANON_LANES) void canon_pairs_kernel(const double* in, size_t m, double* out)
{
    __shared__ double lds[2 * CANON_LANES];
    const size_t base = size_t(blockIdx.x) * CANON_BLOCK;
    // unconditional loads (clamped index, m >= 1) so that all are in flight together
    double va[CANON_PER_LANE], vb[CANON_PER_LANE];
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t i = std::min(base + size_t(j) * CANON_LANES + threadIdx.x, m - 1);
        va[j] = in[2 * i];
        vb[j] = in[2 * i + 1];
    }
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const bool ok = base + size_t(j) * CANON_LANES + threadIdx.x < m;
        a = ok ? a + va[j] : a;
        b = ok ? b + vb[j] : b;
    }
    canon_tree(a, b, lds);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = a;
        out[2 * blockIdx.x + 1] = b;
    }
}

__global__ void fill_f64_kernel(double* p, double v, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

constexpr int SCAN_WG = 1024;

__global__ __launch_bounds__(SCAN_WG) void exclusive_scan_kernel(const uint32_t* counts, int m, uint32_t* offsets, uint32_t* cursor)
{
    __shared__ uint32_t part[SCAN_WG];
    const int t = threadIdx.x;
    const int chunk = (m + SCAN_WG - 1) / SCAN_WG;
    const int b = t * chunk, e = min(m, b + chunk);
    uint32_t s = 0;
    for (int i = b; i < e; i++)
        s += counts[i];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums
    for (int off = 1; off < SCAN_WG; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = b; i < e; i++) {
        offsets[i] = run;
        cursor[i] = run;
        run += counts[i];
    }
    if (t == SCAN_WG - 1)
        offsets[m] = part[t];
}

} // namespace

namespace stp_hip {

const double* canon_reduce_to_block(stp_ctx* ctx, const double* partials, size_t m, double* scratch, size_t* m_out)
{
    const double* src = partials;
    size_t cur = m;
    double* bufs[2] = { scratch, scratch + 2 * ((m + CANON_BLOCK - 1) / CANON_BLOCK) };
    int which = 0;
    while (cur > size_t(CANON_BLOCK)) {
        const size_t nb = (cur + CANON_BLOCK - 1) / CANON_BLOCK;
        double* dst = bufs[which];
        launch(ctx, "canon_reduce", canon_pairs_kernel, dim3(unsigned(nb)), dim3(CANON_LANES), 0, src, cur, dst);
        src = dst;
        cur = nb;
        which ^= 1;
    }
    *m_out = cur;
    return src;
}

void fill_f64(stp_ctx* ctx, double* p, double v, size_t n)
{
    launch(ctx, "fill", fill_f64_kernel, dim3(div_up(n, 256)), dim3(256), 0, p, v, n);
}

void fill_u32(stp_ctx* ctx, uint32_t* p, uint32_t v, size_t n)
{
    launch(ctx, "fill", fill_u32_kernel, dim3(div_up(n, 256)), dim3(256), 0, p, v, n);
}

void exclusive_scan_u32(stp_ctx* ctx, const uint32_t* counts, int m, uint32_t* offsets, uint32_t* cursor)
{
    launch(ctx, "tile_scan", exclusive_scan_kernel, dim3(1), dim3(SCAN_WG), 0, counts, m, offsets, cursor);
}

} // namespace stp_hip

namespace {

constexpr int SCAN_CHUNK = 4096; // 256 threads x 16 consecutive values

struct LoadU32 {
    const uint32_t* p;
    __device__ uint32_t operator()(size_t i) const { return p[i]; }
};
struct LoadPopc64 {
    const unsigned long long* p;
    __device__ uint32_t operator()(size_t i) const { return uint32_t(__popcll(p[i])); }
};

// The 16 values of a thread: straight-line loads when all are in range, so that
// they are in flight together.
template <typename L>
__device__ __forceinline__ void load16(L in, size_t base, size_t m, uint32_t (&v)[16])
{
    if (base + 16 <= m) {
#pragma unroll
        for (int j = 0; j < 16; j++)
            v[j] = in(base + j);
        return;
    }
#pragma unroll
    for (int j = 0; j < 16; j++)
        v[j] = (base + j < m) ? in(base + j) : 0u;
}

// Inclusive scan over the 256 threads of a block: wave shuffles, then the wave totals.
__device__ __forceinline__ uint32_t block_scan256(uint32_t a, uint32_t* wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(a, off, 64);
        if (lane >= off)
            a += y;
    }
    if (lane == 63)
        wsum[w] = a;
    __syncthreads();
    for (int k = 0; k < w; k++)
        a += wsum[k];
    return a;
}

template <typename L>
__global__ __launch_bounds__(256) void chunk_sum_kernel(L in, size_t m, uint32_t* sums)
{
    __shared__ uint32_t wsum[4];
    const size_t base = size_t(blockIdx.x) * SCAN_CHUNK + size_t(threadIdx.x) * 16;
    uint32_t v[16];
    load16(in, base, m, v);
    uint32_t a = 0;
#pragma unroll
    for (int j = 0; j < 16; j++)
        a += v[j];
    a = block_scan256(a, wsum);
    if (threadIdx.x == 255)
        sums[blockIdx.x] = a;
}

// chunk_off: the chunks' offsets; when null, a block adds up the chunk sums before
// it itself (at most RESCAN_SELF chunks: 16 loads per thread, in flight together).
constexpr int RESCAN_SELF = 4096;
template <typename L>
__global__ __launch_bounds__(256) void chunk_rescan_kernel(L in, size_t m, const uint32_t* chunk_off, const uint32_t* sums,
    uint32_t* out)
{
    __shared__ uint32_t wsum[4], wsum2[4];
    __shared__ uint32_t s_off;
    const size_t base = size_t(blockIdx.x) * SCAN_CHUNK + size_t(threadIdx.x) * 16;
    uint32_t v[16];
    load16(in, base, m, v);
    uint32_t off;
    if (chunk_off) {
        off = chunk_off[blockIdx.x];
    } else {
        uint32_t w[RESCAN_SELF / 256];
#pragma unroll
        for (int j = 0; j < RESCAN_SELF / 256; j++) {
            const unsigned c = threadIdx.x + 256u * j;
            w[j] = c < blockIdx.x ? sums[c] : 0u;
        }
        uint32_t p = 0;
#pragma unroll
        for (int j = 0; j < RESCAN_SELF / 256; j++)
            p += w[j];
        p = block_scan256(p, wsum2);
        if (threadIdx.x == 255)
            s_off = p;
        __syncthreads();
        off = s_off;
    }
    uint32_t a = 0;
#pragma unroll
    for (int j = 0; j < 16; j++)
        a += v[j];
    uint32_t run = off + block_scan256(a, wsum) - a;
    for (int j = 0; j < 16; j++) {
        if (base + j < m)
            out[base + j] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255)
        out[m] = run;
}

} // namespace

namespace stp_hip {

template <typename L>
void scan_impl(stp_ctx* ctx, L in, size_t m, uint32_t* out)
{
    const size_t nb = std::max<size_t>(1, (m + SCAN_CHUNK - 1) / SCAN_CHUNK);
    STP_REQUIRE(nb <= (1u << 24), "scan too large");
    uint32_t* sums = (uint32_t*)ws_get(ctx, "scan_sums", sizeof(uint32_t) * nb);
    uint32_t* offs = (uint32_t*)ws_get(ctx, "scan_offs", sizeof(uint32_t) * (nb + 1));
    uint32_t* cur = (uint32_t*)ws_get(ctx, "scan_cur", sizeof(uint32_t) * nb);
    launch(ctx, "scan_sums", chunk_sum_kernel<L>, dim3(unsigned(nb)), dim3(256), 0, in, m, sums);
    if (nb <= size_t(RESCAN_SELF)) {
        // few chunks: every rescan block adds up the sums before it (no scan launch)
        launch(ctx, "scan_rescan", chunk_rescan_kernel<L>, dim3(unsigned(nb)), dim3(256), 0, in, m, (const uint32_t*)nullptr,
            (const uint32_t*)sums, out);
        return;
    }
    exclusive_scan_u32(ctx, sums, int(nb), offs, cur);
    launch(ctx, "scan_rescan", chunk_rescan_kernel<L>, dim3(unsigned(nb)), dim3(256), 0, in, m, (const uint32_t*)offs,
        (const uint32_t*)sums, out);
}

void scan_u32(stp_ctx* ctx, const uint32_t* in, size_t m, uint32_t* out)
{
    scan_impl(ctx, LoadU32{ in }, m, out);
}

void scan_popc64(stp_ctx* ctx, const unsigned long long* in, size_t m, uint32_t* out)
{
    scan_impl(ctx, LoadPopc64{ in }, m, out);
}

} // namespace stp_hip
