// GPU source finder for gfx950: SourceFindImage (src/stp/sourcefind/sourcefind.cpp:156-671).
//
//  statistics  mat_mean_and_stddev (matrix_math.cpp:457-499) with the canonical
//              summation order shared with the oracle (reduce.h), so mean,
//              sigma, the BINAPPROX median and the clipped RMS are bit-identical;
//  median      BINAPPROX: 1001-bin histogram (matrix_math.cpp:334-455);
//              BINMEDIAN / NTHELEMENT: exact order statistics by a linear-bin
//              histogram, candidate gather and an in-LDS bitonic sort, with a
//              radix-select fallback for degenerate data;
//  RMS         estimate_rms sigma clipping (sourcefind.cpp:22-154), device-side
//              iteration state, no host round trips;
//  labelling   union-find over the visual (fft-shifted) image by 64 x 64 tiles
//              (LDS inside a tile, global across tile borders) with the
//              reference's ids: rank of each component's first pixel in the
//              scan (visual column, memory row) order (ccl.h:107-484,505-773);
//  islands     counts, extrema (ties to the smallest memory index) and moments
//              summed per tile component in LDS and combined across tiles,
//              detection validity, label-map cleaning (sourcefind.cpp:342-628);
//              the island list and the moments fit (:235-278, :691-717) are
//              assembled on the device.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "reduce.h"
#include "sourcefind.h"

#pragma clang fp contract(off)

namespace stp_hip {
namespace {

constexpr int SWG = 256;
constexpr int LIN_BINS = 8192;     // exact-median first level
constexpr int SORT_CAP = 8192;     // candidates sorted in one workgroup
constexpr int BA_BINS = 1001;      // BINAPPROX: N + 1 bins (matrix_math.cpp:339,353)

struct SFState {
    double s1, s2;                 // canonical sums
    double mean, sigma, median;
    double rms;
    // sigma-clipping state (sourcefind.cpp:39-151)
    double d1, d2, clip_sigma, upper, prev_upper;
    unsigned long long valid_n;
    int iter, iters, done;
    // device-side control of the median path
    int need_exact;                // exact order statistics still required
    int ba_active;                 // the BINAPPROX histogram is in use
    int compact;                   // the next clipping pass reads the candidates only
    double num_sigma;
    double cand_bound;             // |v - median| bound of the kept clipping candidates
    // thresholds (sourcefind.cpp:298-301)
    double a_pos, a_neg, d_pos, d_neg;
    // results read back by the host
    unsigned int num_pos, num_neg;           // labels of each sign
    unsigned int num_valid, num_islands;     // detections; islands after the area cut
    int fallback;                            // exact median needs the radix selection
    int skip;                                // labelling incomplete (records overflow or median
                                             // fallback): the island stages do nothing
    unsigned int num_records;                // pool component records requested (capacity check)
    unsigned int pad3;
};

template <typename T>
__device__ __forceinline__ double ldv(const T* p, size_t i)
{
    return double(p[i]);
}

// Monotone map of doubles onto unsigned 64-bit keys (-0 and +0 tie).
__device__ __forceinline__ unsigned long long okey(double x)
{
    x = x + 0.0;
    const unsigned long long b = __double_as_longlong(x);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_inv(unsigned long long k)
{
    const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// Visits every element of x once, in no particular order (histograms, gathers):
// 16-byte vector loads, four per thread in flight, grid-stride.  f(double value).
constexpr int HWG = 1024;
template <typename T> struct Vec16;
template <> struct Vec16<float> { using V = float4; static constexpr int W = 4; };
template <> struct Vec16<double> { using V = double2; static constexpr int W = 2; };

template <typename T, typename F>
__device__ __forceinline__ void stream_values(const T* x, size_t n, F&& f)
{
    using V = typename Vec16<T>::V;
    constexpr int W = Vec16<T>::W, U = 4;
    const size_t stride = size_t(gridDim.x) * blockDim.x, gt = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if ((reinterpret_cast<uintptr_t>(x) & 15u) != 0) {
        for (size_t i = gt; i < n; i += stride)
            f(double(x[i]));
        return;
    }
    const size_t nv = n / W;
    const V* xv = reinterpret_cast<const V*>(x);
    size_t i = gt;
    for (; i + (U - 1) * stride < nv; i += U * stride) {
        V v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            v[u] = xv[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const T* e = reinterpret_cast<const T*>(&v[u]);
#pragma unroll
            for (int k = 0; k < W; k++)
                f(double(e[k]));
        }
    }
    for (; i < nv; i += stride) {
        const V v = xv[i];
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int k = 0; k < W; k++)
            f(double(e[k]));
    }
    if (gt < n - nv * W)
        f(double(x[nv * W + gt]));
}

// ------------------------------------------------------------------ sums
// A workgroup takes CANON_GROUP consecutive canonical blocks, loading the next block's
// terms while it reduces the current one.  One block per workgroup measured fastest at
// N = 4096 (sums 17.9 -> 16.2 us, first clipping pass 26.6 -> 21.3 us against four).
constexpr int CANON_GROUP = 1;

template <typename T>
__device__ __forceinline__ void canon_load(const T* x, size_t n, size_t blk, T (&v)[CANON_PER_LANE])
{
    const size_t base = blk * CANON_BLOCK + threadIdx.x;
    if ((blk + 1) * CANON_BLOCK <= n) {
        // a whole block: straight-line loads, all in flight together
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++)
            v[j] = x[base + size_t(j) * CANON_LANES];
        return;
    }
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t i = base + size_t(j) * CANON_LANES;
        v[j] = i < n ? x[i] : T(0);
    }
}

template <typename T>
__global__ __launch_bounds__(SWG) void sf_sum_kernel(const T* x, size_t n, size_t nblk, double* part)
{
    __shared__ double lds[2 * CANON_LANES];
    const size_t b0 = size_t(blockIdx.x) * CANON_GROUP;
    T cur[CANON_PER_LANE], nxt[CANON_PER_LANE];
    canon_load(x, n, b0, cur);
    for (int g = 0; g < CANON_GROUP && b0 + g < nblk; g++) {
        const size_t blk = b0 + g;
        if (g + 1 < CANON_GROUP && blk + 1 < nblk)
            canon_load(x, n, blk + 1, nxt);
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++) {
            const bool ok = blk * CANON_BLOCK + size_t(j) * CANON_LANES + threadIdx.x < n;
            const double v = double(cur[j]);
            a = ok ? a + v : a;
            b = ok ? b + v * v : b;
        }
        canon_tree(a, b, lds);
        if (threadIdx.x == 0) {
            part[2 * blk] = a;
            part[2 * blk + 1] = b;
        }
        __syncthreads();  // lds is rewritten by the next block's tree
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++)
            cur[j] = nxt[j];
    }
}

__device__ void set_mean_sigma(double s1, double s2, unsigned long long n, SFState* st)
{
    const double mean = s1 / double(n);
    st->s1 = s1;
    st->s2 = s2;
    st->mean = mean;
    st->sigma = sqrt(s2 / double(n) - (mean * mean));
}

// ------------------------------------------------------------------ histograms
// bin index of matrix_math.cpp:370-392 with the x86 out-of-range conversion made explicit
__device__ __forceinline__ int ba_bin(double x, double leftend, double scale)
{
    const double t = (x - leftend) * scale;
    if (!(t > -2147483649.0 && t < 2147483648.0))
        return INT_MIN;
    return int(t);
}

template <typename T>
__global__ __launch_bounds__(HWG) void sf_binapprox_hist(const T* x, size_t n, const SFState* st, uint32_t* bh)
{
    __shared__ uint32_t h[BA_BINS + 1];
    if (!st->ba_active)
        return;
    for (int k = threadIdx.x; k < BA_BINS + 1; k += HWG)
        h[k] = 0;
    __syncthreads();
    const double scale = double(BA_BINS - 1) / (2 * st->sigma);
    const double left = st->mean - st->sigma;
    stream_values(x, n, [&](double v) {
        const int b = ba_bin(v, left, scale);
        if (b < 0)
            atomicAdd(&h[0], 1u);
        else if (b <= BA_BINS - 1)
            atomicAdd(&h[1 + b], 1u);
    });
    __syncthreads();
    for (int k = threadIdx.x; k < BA_BINS + 1; k += HWG)
        bh[size_t(blockIdx.x) * (BA_BINS + 1) + k] = h[k];
}

// radix digit histogram (8 bits at `shift`) among elements whose key >> (shift + 8) == prefix
template <typename T>
__global__ __launch_bounds__(HWG) void sf_radix_hist(const T* x, size_t n, int shift, unsigned long long prefix, int use_prefix, uint32_t* bh)
{
    __shared__ uint32_t h[256];
    if (threadIdx.x < 256)
        h[threadIdx.x] = 0;
    __syncthreads();
    stream_values(x, n, [&](double v) {
        const unsigned long long k = okey(v);
        if (!use_prefix || (k >> (shift + 8)) == prefix)
            atomicAdd(&h[(k >> shift) & 0xff], 1u);
    });
    __syncthreads();
    if (threadIdx.x < 256)
        bh[size_t(blockIdx.x) * 256 + threadIdx.x] = h[threadIdx.x];
}

// out[k] += sum over per-block histograms b of bh[b][k]; blockIdx.y takes a
// slice of the blocks, so the launch is (ceil(nb / 256), slices).  out must be
// zero beforehand.  gate: the reduce runs only when *gate != 0 (NULL: always).
__global__ __launch_bounds__(256) void sf_hist_reduce(const uint32_t* bh, int nblk, int nb, unsigned long long* out,
    const int* gate)
{
    if (gate && !*gate)
        return;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nb)
        return;
    const int per = (nblk + int(gridDim.y) - 1) / int(gridDim.y);
    const int b0 = int(blockIdx.y) * per, b1 = min(nblk, b0 + per);
    unsigned long long s = 0;
    for (int c = b0; c < b1; c += 8) {
        uint32_t v[8];  // a batch of loads in flight together (clamped index)
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = bh[size_t(min(c + j, b1 - 1)) * nb + k];
#pragma unroll
        for (int j = 0; j < 8; j++)
            s += c + j < b1 ? v[j] : 0u;
    }
    if (s)
        atomicAdd(&out[k], s);
}

__global__ void sf_ba_gate(SFState* st)
{
    // sourcefind.cpp:187-191: a near-constant image takes the exact median
    const bool exact = (st->sigma * 2) < 0.0001;
    st->ba_active = exact ? 0 : 1;
    st->need_exact = exact ? 1 : 0;
}

// BINAPPROX median from the histogram (matrix_math.cpp:406-452); found = 0 -> exact fallback.
__global__ void sf_binapprox_pick(const unsigned long long* h, unsigned long long n, SFState* st)
{
    if (!st->ba_active)
        return;
    const int N = BA_BINS - 1;
    const double scale = double(N) / (2 * st->sigma);
    const double left = st->mean - st->sigma;
    unsigned long long count = h[0];
    st->need_exact = 1;
    if (n & 1) {
        const unsigned long long k = (n + 1) / 2;
        for (unsigned i = 0; i < unsigned(N + 1); i++) {
            count += h[1 + i];
            if (count >= k) {
                st->median = double(i + 0.5) / scale + left;
                st->need_exact = 0;
                return;
            }
        }
    } else {
        const unsigned long long k = n / 2;
        for (unsigned i = 0; i < unsigned(N + 1); i++) {
            count += h[1 + i];
            if (count >= k) {
                unsigned j = i;
                while (count == k) {
                    j++;
                    count += h[1 + j];
                    if (j == unsigned(N))
                        break;
                }
                st->median = double(i + j + 1) / (2 * scale) + left;
                st->need_exact = 0;
                return;
            }
        }
    }
}

// ------------------------------------------------------------------ candidate gather + sort
struct GatherSpec {
    int shift; // candidates: key >> shift == prefix (shift 64: every element)
    unsigned long long prefix;
};

template <typename T>
__global__ __launch_bounds__(HWG) void sf_gather(const T* x, size_t n, GatherSpec gs, unsigned long long* cand, unsigned* counter, unsigned cap)
{
    stream_values(x, n, [&](double v) {
        const bool take = (gs.shift >= 64) || ((okey(v) >> gs.shift) == gs.prefix);
        if (take) {
            const unsigned slot = atomicAdd(counter, 1u);
            if (slot < cap)
                cand[slot] = okey(v);
        }
    });
}

// Ascending bitonic sort of the first c (<= SORT_CAP) keys of cand into s
// (1024 threads).
__device__ void lds_sort_keys(const unsigned long long* cand, unsigned c, unsigned long long* s)
{
    unsigned m = 1;
    while (m < c)
        m <<= 1;
    for (unsigned i = threadIdx.x; i < m; i += 1024)
        s[i] = (i < c) ? cand[i] : ~0ull;
    __syncthreads();
    for (unsigned k = 2; k <= m; k <<= 1) {
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = threadIdx.x; i < m; i += 1024) {
                const unsigned l = i ^ j;
                if (l > i) {
                    const unsigned long long a = s[i], b = s[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Sorts count (<= SORT_CAP) keys in LDS and writes the values of ranks r0 and r1
// (0-based within the candidates) to out[0], out[1].
__global__ __launch_bounds__(1024) void sf_sort_select(const unsigned long long* cand, const unsigned* count, int r0, int r1, double* out)
{
    __shared__ unsigned long long s[SORT_CAP];
    const unsigned c = min(*count, unsigned(SORT_CAP));
    lds_sort_keys(cand, c, s);
    if (threadIdx.x == 0) {
        out[0] = (r0 >= 0 && unsigned(r0) < c) ? okey_inv(s[r0]) : 0.0;
        out[1] = (r1 >= 0 && unsigned(r1) < c) ? okey_inv(s[r1]) : 0.0;
    }
}

__global__ void sf_set_median(SFState* st, const double* lohi, int odd)
{
    // mat_median_exact (matrix_math.cpp:33-45): (upper + lower) / 2 for even sizes
    st->median = odd ? lohi[1] : (lohi[1] + lohi[0]) / 2;
}

// ------------------------------------------------------------------ exact selection
// Device-resident search for the order statistics x_(k_lo) <= x_(k_hi): level 1
// bins mean +- 1.0625 sigma linearly (the median lies within one sigma of the
// mean); when the level-1 bins holding both ranks contain more than SORT_CAP
// elements, level 2 splits those bins into LIN_BINS sub-bins.  The candidates of
// the final bins are gathered and sorted in LDS.  Every kernel reads its gate
// from SelState, so the whole search is enqueued without host round trips;
// `fallback` asks the host for a radix selection (heavily tied data).
struct SelState {
    double lo, scale;               // level 1: t = (v - lo) * scale
    double base2, scale2;           // level 2: t2 = (t - base2) * scale2
    unsigned long long k_lo, k_hi;  // ranks sought (0-based)
    unsigned long long before;      // elements ranked below the candidate set
    unsigned long long cnt;         // elements in the candidate set
    int run1, run2, gather, fallback;
    int b1_lo, b1_hi, b2_lo, b2_hi;
    int level, odd;
};

__device__ __forceinline__ int lin_bin(double t)
{
    return t < 0.0 ? 0 : (t < double(LIN_BINS) ? 1 + int(t) : LIN_BINS + 1);
}
__device__ __forceinline__ int sub_bin(double t, double base2, double scale2)
{
    const int b = int((t - base2) * scale2);
    return b < 0 ? 0 : (b >= LIN_BINS ? LIN_BINS - 1 : b);
}

// Bins holding ranks ka <= kb of a histogram of nb <= 1024 * RB_PER bins (1024
// threads).  Writes the bin indices, the count before bin ia and the count
// through bin ib; -1 if absent.
constexpr int RB_PER = 9;
static_assert(LIN_BINS + 2 <= 1024 * RB_PER, "rank_bins capacity");
__device__ void rank_bins(const unsigned long long* h, int nb, unsigned long long ka, unsigned long long kb, int* ia,
    int* ib, unsigned long long* before_a, unsigned long long* through_b)
{
    __shared__ unsigned long long wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) {
        *ia = -1;
        *ib = -1;
    }
    const int per = (nb + 1023) / 1024;
    const int i0 = min(nb, t * per), i1 = min(nb, i0 + per);
    // the thread's bins, all loads in flight together (clamped index)
    unsigned long long v[RB_PER];
#pragma unroll
    for (int j = 0; j < RB_PER; j++)
        v[j] = h[min(i0 + j, nb - 1)];
    unsigned long long s = 0;
#pragma unroll
    for (int j = 0; j < RB_PER; j++)
        s += i0 + j < i1 ? v[j] : 0ull;
    // exclusive scan over the threads: wave shuffles, then the wave totals
    unsigned long long x = s;
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    unsigned long long cum = x - s;
    for (int k = 0; k < w; k++)
        cum += wsum[k];
#pragma unroll
    for (int j = 0; j < RB_PER; j++) {
        if (i0 + j >= i1)
            break;
        const unsigned long long c = cum + v[j];
        if (cum <= ka && c > ka) {
            *ia = i0 + j;
            *before_a = cum;
        }
        if (cum <= kb && c > kb) {
            *ib = i0 + j;
            *through_b = c;
        }
        cum = c;
    }
    __syncthreads();
}

struct SelInit {
    int enabled;                     // run the selection set-up in the statistics kernel
    unsigned long long k_lo, k_hi;
    int odd, force;
    unsigned long long* hist1;       // LIN_BINS + 2 level-1 counts
    unsigned long long* hist2;       // LIN_BINS level-2 counts
    unsigned* counter;               // gathered candidates
};

// Selection set-up (thread 0) and cleared histograms and candidate counter (the block).
__device__ void sel_init_block(SFState* st, SelState* ss, const SelInit& si)
{
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += blockDim.x)
        si.hist1[k] = 0;
    for (int k = threadIdx.x; k < LIN_BINS; k += blockDim.x)
        si.hist2[k] = 0;
    if (threadIdx.x != 0)
        return;
    *si.counter = 0;
    const unsigned long long k_lo = si.k_lo, k_hi = si.k_hi;
    const int odd = si.odd, force = si.force;
    if (force)
        st->need_exact = 1;
    const bool active = st->need_exact != 0;
    const double m = st->mean, sg = st->sigma;
    const bool binnable = sg > 0.0 && isfinite(sg) && isfinite(m);
    ss->k_lo = k_lo;
    ss->k_hi = k_hi;
    ss->odd = odd;
    ss->before = 0;
    ss->cnt = 0;
    ss->run1 = (active && binnable) ? 1 : 0;
    ss->run2 = 0;
    ss->gather = 0;
    ss->fallback = (active && !binnable) ? 1 : 0;
    ss->level = 0;
    const double lo = m - 1.0625 * sg, hi = m + 1.0625 * sg;
    ss->lo = lo;
    ss->scale = double(LIN_BINS) / (hi - lo);
}

__global__ __launch_bounds__(256) void sel_init(SFState* st, SelState* ss, SelInit si)
{
    sel_init_block(st, ss, si);
}

// Mean and sigma from the block partials of sf_sum (last canonical level), then
// optionally the selection set-up.  clear (nclear doubles, may be null): the state
// buffer, zeroed first.
__global__ __launch_bounds__(CANON_LANES) void sf_stats(const double* part, unsigned m, unsigned long long n, SFState* st,
    SelState* ss, SelInit si, double* clear, unsigned nclear)
{
    __shared__ double lds[2 * CANON_LANES];
    if (clear)
        for (unsigned k = threadIdx.x; k < nclear; k += CANON_LANES)
            clear[k] = 0.0;
    double a, b;
    canon_last_level(part, m, a, b, lds);  // its barriers also order the clearing before the writes
    if (threadIdx.x == 0)
        set_mean_sigma(a, b, n, st);
    if (si.enabled)
        sel_init_block(st, ss, si);
}

template <typename T>
__global__ __launch_bounds__(HWG) void sel_hist1(const T* x, size_t n, const SelState* ss, uint32_t* bh)
{
    __shared__ uint32_t h[LIN_BINS + 2];
    if (!ss->run1)
        return;
    const double lo = ss->lo, scale = ss->scale;
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += HWG)
        h[k] = 0;
    __syncthreads();
    stream_values(x, n, [&](double v) { atomicAdd(&h[lin_bin((v - lo) * scale)], 1u); });
    __syncthreads();
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += HWG)
        bh[size_t(blockIdx.x) * (LIN_BINS + 2) + k] = h[k];
}

__global__ __launch_bounds__(1024) void sel_pick1(SelState* ss, const unsigned long long* hist)
{
    __shared__ int ia, ib;
    __shared__ unsigned long long ba, tb;
    if (!ss->run1)
        return;
    rank_bins(hist, LIN_BINS + 2, ss->k_lo, ss->k_hi, &ia, &ib, &ba, &tb);
    if (threadIdx.x != 0)
        return;
    // both ranks must fall inside the binned range (not in the below/above bins)
    if (ia < 1 || ib < ia || ib > LIN_BINS) {
        ss->fallback = 1;
        return;
    }
    ss->b1_lo = ia;
    ss->b1_hi = ib;
    ss->before = ba;
    ss->cnt = tb - ba;
    if (tb - ba <= (unsigned long long)SORT_CAP) {
        ss->level = 1;
        ss->gather = 1;
    } else {
        ss->base2 = double(ia - 1);
        ss->scale2 = double(LIN_BINS) / double(ib - ia + 1);
        ss->run2 = 1;
    }
}

template <typename T>
__global__ __launch_bounds__(HWG) void sel_hist2(const T* x, size_t n, const SelState* ss, uint32_t* bh)
{
    __shared__ uint32_t h[LIN_BINS];
    if (!ss->run2)
        return;
    const double lo = ss->lo, scale = ss->scale, base2 = ss->base2, scale2 = ss->scale2;
    const int b1lo = ss->b1_lo, b1hi = ss->b1_hi;
    for (int k = threadIdx.x; k < LIN_BINS; k += HWG)
        h[k] = 0;
    __syncthreads();
    stream_values(x, n, [&](double v) {
        const double t = (v - lo) * scale;
        const int b = lin_bin(t);
        if (b >= b1lo && b <= b1hi)
            atomicAdd(&h[sub_bin(t, base2, scale2)], 1u);
    });
    __syncthreads();
    for (int k = threadIdx.x; k < LIN_BINS; k += HWG)
        bh[size_t(blockIdx.x) * LIN_BINS + k] = h[k];
}

__global__ __launch_bounds__(1024) void sel_pick2(SelState* ss, const unsigned long long* hist)
{
    __shared__ int ia, ib;
    __shared__ unsigned long long ba, tb;
    if (!ss->run2)
        return;
    const unsigned long long b0 = ss->before;
    rank_bins(hist, LIN_BINS, ss->k_lo - b0, ss->k_hi - b0, &ia, &ib, &ba, &tb);
    if (threadIdx.x != 0)
        return;
    if (ia < 0 || ib < ia || tb - ba > (unsigned long long)SORT_CAP) {
        ss->fallback = 1;
        return;
    }
    ss->b2_lo = ia;
    ss->b2_hi = ib;
    ss->before = b0 + ba;
    ss->cnt = tb - ba;
    ss->level = 2;
    ss->gather = 1;
}

template <typename T>
__global__ __launch_bounds__(HWG) void sel_gather(const T* x, size_t n, const SelState* ss, unsigned long long* cand,
    unsigned* counter)
{
    if (!ss->gather)
        return;
    const double lo = ss->lo, scale = ss->scale, base2 = ss->base2, scale2 = ss->scale2;
    const int b1lo = ss->b1_lo, b1hi = ss->b1_hi, b2lo = ss->b2_lo, b2hi = ss->b2_hi, lvl = ss->level;
    stream_values(x, n, [&](double v) {
        const double t = (v - lo) * scale;
        const int b = lin_bin(t);
        bool take = b >= b1lo && b <= b1hi;
        if (take && lvl == 2) {
            const int b2 = sub_bin(t, base2, scale2);
            take = b2 >= b2lo && b2 <= b2hi;
        }
        if (take) {
            const unsigned slot = atomicAdd(counter, 1u);
            if (slot < unsigned(SORT_CAP))
                cand[slot] = okey(v);
        }
    });
}

// Sorts the candidates and sets the median (mat_median_exact, matrix_math.cpp:33-45:
// (upper + lower) / 2 for even sizes).
__global__ __launch_bounds__(1024) void sel_sort(SelState* ss, const unsigned long long* cand, const unsigned* count,
    SFState* st)
{
    __shared__ unsigned long long s[SORT_CAP];
    if (!ss->gather)
        return;
    const unsigned c = min(*count, unsigned(SORT_CAP));
    lds_sort_keys(cand, c, s);
    if (threadIdx.x != 0)
        return;
    const unsigned long long r0 = ss->k_lo - ss->before, r1 = ss->k_hi - ss->before;
    if (c != ss->cnt || r1 >= c) {
        ss->fallback = 1; // the gather disagrees with the histogram counts
        return;
    }
    const double lo = okey_inv(s[r0]), hi = okey_inv(s[r1]);
    st->median = ss->odd ? hi : (hi + lo) / 2;
    st->need_exact = 0;
}

// ------------------------------------------------------------------ sigma clipping
// A pass over the whole image also keeps, per canonical block and lane in the
// lane's order, the values with |v - median| > CLIP_CAND_FRAC * upper (the
// candidates).  While the clipping bound stays at or above that value, the next
// passes read only the candidates: every other term of their sums is an exact
// zero, so the sums equal those of a full pass bit for bit.  A pass whose bound
// has dropped below runs over the whole image again and keeps new candidates.
// 0.5: on dirty images of noise and point sources the bound falls by 10-25 % a pass
// and the candidates are ~8 % of the pixels.
constexpr double CLIP_CAND_FRAC = 0.5;

// Removal of one term (sourcefind.cpp:39-151): its share of the sums of the removed.
__device__ __forceinline__ void clip_term(double x, double median, double upper, double prev, double& a, double& b,
    unsigned& cnt)
{
    const double v = x - median;
    const double av = fabs(v);
    const bool rem = (av > upper) && !(av > prev);
    a = a + (rem ? v : 0.0);
    b = b + (rem ? v * v : 0.0);
    cnt += rem ? 1u : 0u;
}

// A candidate pass over one canonical block by one wave: lane l stands for the
// canonical lanes l + 64 q (q < 4).  The levels of canon_tree follow: lanes t and
// t + 128, then t and t + 64, then the shuffles of the wave.  The first CLIP_BATCH
// candidates of every lane are loaded together; the rest (rare) in a second batch.
constexpr int CLIP_BATCH = 8;
template <typename T>
__device__ __forceinline__ void clip_candidates(const T* cand, const uint32_t* ltab, size_t blk, double median,
    double upper, double prev, double* part, uint32_t* removed)
{
    const int lane = threadIdx.x & 63;
    uint32_t lw[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
        lw[q] = ltab[blk * CANON_LANES + 64 * q + lane];
    T v[4][CLIP_BATCH];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const T* cp = cand + blk * CANON_BLOCK + (lw[q] & 0xFFFFu);
        const int c = int(lw[q] >> 16);
#pragma unroll
        for (int j = 0; j < CLIP_BATCH; j++)
            v[q][j] = j < c ? cp[j] : T(0);
    }
    double a[4], b[4];
    unsigned cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        a[q] = 0.0;
        b[q] = 0.0;
        const int c = int(lw[q] >> 16);
#pragma unroll
        for (int j = 0; j < CLIP_BATCH; j++)
            if (j < c)
                clip_term(double(v[q][j]), median, upper, prev, a[q], b[q], cnt);
    }
    const uint32_t cmax = max(max(lw[0], lw[1]), max(lw[2], lw[3])) >> 16;
    if (__any(cmax > CLIP_BATCH)) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const T* cp = cand + blk * CANON_BLOCK + (lw[q] & 0xFFFFu);
            const int c = int(lw[q] >> 16);
#pragma unroll
            for (int j = CLIP_BATCH; j < CANON_PER_LANE; j++)
                v[q][j - CLIP_BATCH] = j < c ? cp[j] : T(0);
#pragma unroll
            for (int j = CLIP_BATCH; j < CANON_PER_LANE; j++)
                if (j < c)
                    clip_term(double(v[q][j - CLIP_BATCH]), median, upper, prev, a[q], b[q], cnt);
        }
    }
    double x = (a[0] + a[2]) + (a[1] + a[3]);
    double y = (b[0] + b[2]) + (b[1] + b[3]);
    for (int off = 32; off >= 1; off >>= 1) {
        x = x + __shfl_down(x, off, 64);
        y = y + __shfl_down(y, off, 64);
        cnt += __shfl_down(cnt, off, 64);
    }
    if (lane == 0) {
        part[2 * blk] = x;
        part[2 * blk + 1] = y;
        removed[blk] = cnt;
    }
}

// A pass over one canonical block by one wave (lanes as in clip_candidates) that
// reads the whole block and keeps its candidates: the pass after a drop of the
// bound below that of the kept candidates.
template <typename T>
__device__ __forceinline__ void clip_block_full(const T* x, size_t n, size_t blk, double median, double upper,
    double prev, double bound, double* part, uint32_t* removed, T* cand, uint32_t* ltab)
{
    const int lane = threadIdx.x & 63;
    double a[4], b[4];
    unsigned cnt = 0;
    uint32_t base = 0;  // candidates of the canonical lanes before 64 q
    for (int q = 0; q < 4; q++) {
        const size_t i0 = blk * CANON_BLOCK + 64 * q + lane;
        T v[CANON_PER_LANE];
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++)
            v[j] = i0 + size_t(j) * CANON_LANES < n ? x[i0 + size_t(j) * CANON_LANES] : T(0);
        double aq = 0.0, bq = 0.0;
        uint32_t keep = 0;
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++) {
            if (i0 + size_t(j) * CANON_LANES < n) {
                clip_term(double(v[j]), median, upper, prev, aq, bq, cnt);
                keep |= (fabs(double(v[j]) - median) > bound) ? (1u << j) : 0u;
            }
        }
        a[q] = aq;
        b[q] = bq;
        const uint32_t kc = uint32_t(__popc(keep));
        uint32_t kx = kc;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(kx, off, 64);
            if (lane >= off)
                kx += y;
        }
        const uint32_t off = base + kx - kc;
        ltab[blk * CANON_LANES + 64 * q + lane] = off | (kc << 16);
        T* cp = cand + blk * CANON_BLOCK + off;
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++) {
            if (keep & (1u << j))
                *cp++ = v[j];
        }
        base += __shfl(kx, 63, 64);
    }
    double xs = (a[0] + a[2]) + (a[1] + a[3]);
    double ys = (b[0] + b[2]) + (b[1] + b[3]);
    for (int off = 32; off >= 1; off >>= 1) {
        xs = xs + __shfl_down(xs, off, 64);
        ys = ys + __shfl_down(ys, off, 64);
        cnt += __shfl_down(cnt, off, 64);
    }
    if (lane == 0) {
        part[2 * blk] = xs;
        part[2 * blk + 1] = ys;
        removed[blk] = cnt;
    }
}

// Clipping passes after the first: one wave per canonical block, reading the kept
// candidates while the bound allows it (every lane of the grid takes one branch).
template <typename T>
__global__ __launch_bounds__(SWG) void sf_clip_next_kernel(const T* x, size_t n, size_t nblk, const SFState* st,
    double* part, uint32_t* removed, T* cand, uint32_t* ltab)
{
    if (st->done)
        return;
    const double upper = st->upper, prev = st->prev_upper, median = st->median;
    const size_t blk = size_t(blockIdx.x) * (SWG / 64) + (threadIdx.x >> 6);
    if (blk >= nblk)
        return;
    if (st->compact)
        clip_candidates(cand, ltab, blk, median, upper, prev, part, removed);
    else
        clip_block_full(x, n, blk, median, upper, prev, CLIP_CAND_FRAC * upper, part, removed, cand, ltab);
}

// The first clipping pass (the state of sf_clip_init) over the whole image; keeps
// the candidates of its bound.
template <typename T>
__global__ __launch_bounds__(SWG) void sf_clip_kernel(const T* x, size_t n, size_t nblk, const SFState* st,
    double num_sigma, double* part, uint32_t* removed, T* cand, uint32_t* ltab)
{
    __shared__ double lds[2 * CANON_LANES];
    __shared__ uint32_t wc[SWG / 64];
    __shared__ uint32_t wk[SWG / 64];
    const double upper = num_sigma * st->sigma, prev = DBL_MAX;
    const double median = st->median;
    const double bound = CLIP_CAND_FRAC * upper;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const size_t b0 = size_t(blockIdx.x) * CANON_GROUP;
    T cur[CANON_PER_LANE], nxt[CANON_PER_LANE];
    canon_load(x, n, b0, cur);
    for (int g = 0; g < CANON_GROUP && b0 + g < nblk; g++) {
        const size_t blk = b0 + g;
        if (g + 1 < CANON_GROUP && blk + 1 < nblk)
            canon_load(x, n, blk + 1, nxt);
        const size_t base = blk * CANON_BLOCK + t;
        const int nv = base >= n ? 0 : int(std::min<size_t>(CANON_PER_LANE, (n - base + CANON_LANES - 1) / CANON_LANES));
        double a = 0.0, b = 0.0;
        unsigned cnt = 0, keep = 0;
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++) {
            if (j < nv) {
                clip_term(double(cur[j]), median, upper, prev, a, b, cnt);
                keep |= (fabs(double(cur[j]) - median) > bound) ? (1u << j) : 0u;
            }
        }
        canon_tree(a, b, lds);
        for (int off = 32; off >= 1; off >>= 1)
            cnt += __shfl_down(cnt, off, 64);
        // exclusive scan of the lanes' candidate counts
        const uint32_t kc = uint32_t(__popc(keep));
        uint32_t kx = kc;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(kx, off, 64);
            if (lane >= off)
                kx += y;
        }
        if (lane == 0)
            wc[w] = cnt;
        if (lane == 63)
            wk[w] = kx;
        __syncthreads();
        if (t == 0) {
            part[2 * blk] = a;
            part[2 * blk + 1] = b;
            uint32_t c = 0;
            for (int k = 0; k < SWG / 64; k++)
                c += wc[k];
            removed[blk] = c;
        }
        uint32_t off = kx - kc;
        for (int k = 0; k < w; k++)
            off += wk[k];
        ltab[blk * CANON_LANES + t] = off | (kc << 16);
        T* cp = cand + blk * CANON_BLOCK + off;
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++) {
            if (keep & (1u << j))
                *cp++ = cur[j];
        }
        __syncthreads();  // lds, wc and wk are rewritten by the next block
#pragma unroll
        for (int j = 0; j < CANON_PER_LANE; j++)
            cur[j] = nxt[j];
    }
}

__device__ void clip_init_state(SFState* st, unsigned long long n, int iters, double num_sigma)
{
    const double median = st->median;
    const double mean = st->mean - median;
    const double sigma = st->sigma;
    st->d1 = double(mean) * double(n);
    st->d2 = (double(sigma) * double(sigma) + double(mean) * double(mean)) * double(n);
    st->valid_n = n;
    st->clip_sigma = sigma;
    st->num_sigma = num_sigma;
    st->upper = num_sigma * sigma;
    st->prev_upper = DBL_MAX;
    st->iter = 0;
    st->iters = iters;
    st->done = iters <= 0;
    st->compact = 0;
}

__global__ void sf_clip_init(SFState* st, unsigned long long n, int iters, double num_sigma)
{
    clip_init_state(st, n, iters, num_sigma);
}

// Detection and analysis thresholds (sourcefind.cpp:298-301).
struct ThreshArgs {
    int estimate;        // RMS from the clipping
    double rms_in, det, ana;
    uint32_t* reset;     // counter cleared for the labelling (may be null)
};

__device__ void set_thresholds(SFState* st, const ThreshArgs& ta)
{
    const double rms = ta.estimate ? st->clip_sigma : ta.rms_in;
    st->rms = rms;
    const double bg = st->median;
    st->a_pos = bg + ta.ana * rms;
    st->a_neg = bg - ta.ana * rms;
    st->d_pos = bg + ta.det * rms;
    st->d_neg = bg - ta.det * rms;
    if (ta.reset)
        *ta.reset = 0;
}

__global__ void sf_thresholds(SFState* st, ThreshArgs ta)
{
    set_thresholds(st, ta);
}

// One clipping step (sourcefind.cpp:39-151) from the block partials of sf_clip
// (m <= CANON_BLOCK pairs after any earlier canonical levels; removed: one count
// per block).  Iteration 0 first sets up the state; the last iteration also sets
// the thresholds.  The partials are read before the state is: after the end of
// the clipping they are stale and go unused.
__global__ __launch_bounds__(256) void sf_clip_update(SFState* st, const double* cpart, unsigned m, const uint32_t* removed,
    int nblk, int iter, unsigned long long n, int iters, double num_sigma, int last, ThreshArgs ta)
{
    __shared__ double lds[2 * CANON_LANES];
    __shared__ unsigned long long red[256 / 64];
    const int t = threadIdx.x;
    unsigned long long c = 0;
    for (int k0 = 0; k0 < nblk; k0 += 256 * 16) {
        // unconditional loads (clamped index) so that none waits on another
        uint32_t r[16];
#pragma unroll
        for (int j = 0; j < 16; j++)
            r[j] = removed[min(k0 + j * 256 + t, nblk - 1)];
#pragma unroll
        for (int j = 0; j < 16; j++)
            c += k0 + j * 256 + t < nblk ? r[j] : 0u;
    }
    double a, b;
    canon_last_level(cpart, m, a, b, lds);
    for (int off = 32; off >= 1; off >>= 1)
        c += __shfl_down(c, off, 64);
    if ((t & 63) == 0)
        red[t >> 6] = c;
    __syncthreads();
    if (t != 0)
        return;
    if (iter == 0)
        clip_init_state(st, n, iters, num_sigma);
    if (!st->done) {
        const unsigned long long rem = red[0] + red[1] + red[2] + red[3];
        // a full pass kept the candidates of its own bound
        if (!st->compact)
            st->cand_bound = CLIP_CAND_FRAC * st->upper;
        if (rem == 0) {
            st->done = 1;
        } else {
            st->valid_n -= rem;
            st->d1 -= a;
            st->d2 -= b;
            const double vn = double(st->valid_n);
            st->clip_sigma = sqrt(st->d2 / vn - (st->d1 / vn) * (st->d1 / vn));
            st->prev_upper = st->upper;
            st->upper = st->num_sigma * st->clip_sigma;
            st->compact = st->upper >= st->cand_bound;
            st->iter++;
            if (st->iter >= st->iters)
                st->done = 1;
        }
    }
    if (last)
        set_thresholds(st, ta);
}

// ------------------------------------------------------------------ labelling
// Connected components (ccl.h:107-484) by tiles.  The visual (fft-shifted)
// image is cut into tiles of CT_W visual columns x CT_H visual rows; tile rows
// are aligned so that no tile straddles visual row R/2, where the memory row
// wraps, so inside a tile the key (visual column * R + memory row) grows with
// the local index c * CT_H + rr.  A component's root is its smallest key, i.e.
// its first pixel in the reference's scan order (visual column, memory row).
//
//  ccl_tile        classification; vertical runs from wave ballots (a wave holds
//                  one tile column); a key-ordered list of the foreground pixels
//                  in LDS, over which the rest runs: union-find of the runs, one
//                  union per pair of touching runs; the island sums of every
//                  component of the tile (count, extremum, its smallest memory
//                  index, moments) in LDS; a component with a same-class
//                  neighbour outside the tile is "open";
//  ccl_border      union-find across tile borders, on the open components only;
//  ccl_open_roots  flattens open components, marks their global roots;
//  ccl_records     one record per tile component: closed ones are final, open
//                  ones are combined into their global root with atomics.
// Label ids are ranks of roots in key order: a 64-bit root mask per segment (one
// tile's rows of one column) and a scan of the segment counts give every root's
// id without a pass over the image.
constexpr int CT_W = 64, CT_H = 64, CT_PIX = CT_W * CT_H, CT_THREADS = 1024, CT_WAVES = CT_THREADS / 64;
constexpr int CT_PPT = CT_PIX / CT_THREADS;  // pixels per thread: columns w + CT_WAVES j, row = lane
constexpr int CT_RCAP = 480;  // components of a tile summed in LDS; more go to global atomics
constexpr uint32_t P_BG = 0xFFFFFFFFu, P_OPEN = 0x80000000u, P_NEG = 0x40000000u, P_KEY = 0x3FFFFFFFu;
constexpr uint32_t L_KEY = 0xFFFu, L_OPEN = 0x1000u;

struct TileGeom {
    int R, C, half_r, half_c;
    int base;    // visual row where tile row 0 would start (<= 0): tile rows are aligned to R/2
    int nty, ntx;
    int ty_mid;  // tile row starting at visual row R/2: key order starts there
    unsigned nseg;
};

TileGeom tile_geom(int rows, int cols)
{
    TileGeom g{};
    g.R = rows;
    g.C = cols;
    g.half_r = rows / 2;
    g.half_c = cols / 2;
    const int off = g.half_r % CT_H;
    g.base = off ? off - CT_H : 0;
    g.nty = (rows - g.base + CT_H - 1) / CT_H;
    g.ntx = (cols + CT_W - 1) / CT_W;
    g.ty_mid = (g.half_r - g.base) / CT_H;
    g.nseg = unsigned(cols) * unsigned(g.nty);
    return g;
}

__device__ __forceinline__ int tile_vr0(const TileGeom& g, int ty) { return max(0, g.base + CT_H * ty); }
__device__ __forceinline__ int tile_vr1(const TileGeom& g, int ty) { return min(g.R, g.base + CT_H * (ty + 1)); }
// visual row <-> memory row (the same map both ways, R even)
__device__ __forceinline__ int flip_row(const TileGeom& g, int r) { return r >= g.half_r ? r - g.half_r : r + g.half_r; }
__device__ __forceinline__ int mem_col(const TileGeom& g, int cv) { return cv >= g.half_c ? cv - g.half_c : cv + g.half_c; }
// segment (visual column cv, tile row ty) in key order
__device__ __forceinline__ unsigned seg_index(const TileGeom& g, int cv, int ty)
{
    const int pos = ty >= g.ty_mid ? ty - g.ty_mid : ty + (g.nty - g.ty_mid);
    return unsigned(cv) * unsigned(g.nty) + unsigned(pos);
}

// Sums of one island (or of one tile component of it).
struct IslandSum {
    unsigned long long ext;  // order key of the extremum
    uint32_t idx;            // smallest memory index holding the extremum
    int cnt;
    double mom[6];           // sum x v, y v, x x v, y y v, x y v, v
};

struct SegIds {
    TileGeom g;
    const unsigned long long* mask;  // roots per segment: [0, nseg) positive, [nseg, 2 nseg) negative;
                                     // then [2 nseg, 3 nseg): the foreground pixels of each segment
    const uint32_t* offs;            // exclusive scan of the popcounts of mask, 2 nseg + 1 entries
};

// Position of a root in the island list: positive ids 0..kp-1, then negative ones.
__device__ __forceinline__ uint32_t gid_at(const SegIds& s, unsigned sseg, int rr)
{
    return s.offs[sseg] + uint32_t(__popcll(s.mask[sseg] & ((1ull << rr) - 1ull)));
}
__device__ __forceinline__ unsigned seg_of_key(const TileGeom& g, uint32_t key, bool neg, int& rr)
{
    const int cv = int(key / unsigned(g.R)), vr = flip_row(g, int(key % unsigned(g.R)));
    const int ty = (vr - g.base) / CT_H;
    rr = vr - tile_vr0(g, ty);
    return seg_index(g, cv, ty) + (neg ? g.nseg : 0u);
}
__device__ __forceinline__ uint32_t root_gid(const SegIds& s, uint32_t key, bool neg)
{
    int rr;
    const unsigned sseg = seg_of_key(s.g, key, neg, rr);
    return gid_at(s, sseg, rr);
}

// Label g's sums: lrec[g] is the record of a closed component, or LREC_OPEN for an
// open island, whose combined sums are isum[g].
constexpr uint32_t LREC_OPEN = 0xFFFFFFFFu;
__device__ __forceinline__ const IslandSum& label_sum(const IslandSum* isum, const uint32_t* lrec, const IslandSum* rsum,
    uint32_t g)
{
    const uint32_t r = lrec[g];
    return r == LREC_OPEN ? isum[g] : rsum[r];
}

// Per-component records written by ccl_tile: tile b owns records [b rt, b rt + rt);
// components beyond rt take records of a shared pool after the tile regions (one
// atomic per overflowing tile: a device-wide counter per tile would serialise).
struct CompRecords {
    uint32_t* key;          // local root key | P_NEG | P_OPEN
    uint32_t* pos;          // (signed segment << 6) | row of the local root in it
    IslandSum* sum;
    uint32_t* tcount;       // components of each tile
    uint32_t* pool_count;   // pool records requested
    uint32_t rt, ntiles, pool_cap;
};

constexpr uint32_t CT_RT = 256;  // records of a tile's own region

// Record i of the tile regions or the pool, when it holds a component.
__device__ __forceinline__ bool rec_valid(const CompRecords& rec, uint32_t i)
{
    const uint32_t reg = rec.ntiles * rec.rt;
    if (i < reg)
        return (i % rec.rt) < min(rec.tcount[i / rec.rt], rec.rt);
    return i - reg < *rec.pool_count;
}
// An overflowed pool leaves records unwritten: the host grows it and labels again.
__device__ __forceinline__ bool pool_overflow(const CompRecords& rec) { return *rec.pool_count > rec.pool_cap; }

// Plain LDS loads: every load has a new address, and the atomics of lds_union
// order the reloads after a failed link.
__device__ __forceinline__ uint32_t lds_find(const uint32_t* par, uint32_t a)
{
    uint32_t p = par[a];
    while (p != a) {
        a = p;
        p = par[a];
    }
    return a;
}

// Links the trees of a and b in LDS; the smaller local index (smaller key) becomes the root.
__device__ void lds_union(uint32_t* par, uint32_t a, uint32_t b)
{
    bool done;
    do {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a < b) {
            const uint32_t old = atomicMin(&par[b], a);
            done = (old == b);
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(&par[a], b);
            done = (old == a);
            a = old;
        } else {
            done = true;
        }
    } while (!done);
}

__device__ __forceinline__ uint32_t g_find(const uint32_t* P, uint32_t a)
{
    const volatile uint32_t* vp = P;
    uint32_t p = vp[a] & P_KEY;
    while (p != a) {
        a = p;
        p = vp[a] & P_KEY;
    }
    return a;
}

// Global union of two open components of one class (flags: P_OPEN and the class bit).
__device__ void g_union(uint32_t* P, uint32_t a, uint32_t b, uint32_t flags)
{
    bool done;
    do {
        a = g_find(P, a);
        b = g_find(P, b);
        if (a < b) {
            const uint32_t old = atomicMin(&P[b], a | flags);
            done = (old == (b | flags));
            b = old & P_KEY;
        } else if (b < a) {
            const uint32_t old = atomicMin(&P[a], b | flags);
            done = (old == (a | flags));
            a = old & P_KEY;
        } else {
            done = true;
        }
    } while (!done);
}

template <typename T>
__global__ __launch_bounds__(CT_THREADS) void ccl_tile(const T* __restrict__ x, TileGeom g, const SFState* st, int find_neg,
    int conn4, uint32_t* __restrict__ P, unsigned long long* __restrict__ rmask, CompRecords rec)
{
    __shared__ uint32_t s_par[CT_PIX];
    __shared__ uint32_t s_list[CT_PIX];  // foreground pixels in key order: li | class << 12 | run start << 14 | run end << 20
    __shared__ uint8_t s_cls[CT_PIX];
    __shared__ uint8_t s_fl[CT_H + 2], s_fr[CT_H + 2], s_ft[CT_W], s_fb[CT_W];
    __shared__ unsigned long long s_mask[CT_W];        // foreground of each column
    __shared__ unsigned long long s_rmp[CT_W], s_rmn[CT_W];  // closed roots of each column
    __shared__ uint32_t s_pre[CT_W + 1];
    __shared__ uint32_t s_base, s_nroot;
    __shared__ unsigned long long r_ext[CT_RCAP];
    __shared__ uint32_t r_idx[CT_RCAP];
    __shared__ int r_cnt[CT_RCAP];
    __shared__ double r_mom[6][CT_RCAP];

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tx = int(blockIdx.x) % g.ntx, ty = int(blockIdx.x) / g.ntx;
    const int cv0 = tx * CT_W, nc = min(CT_W, g.C - cv0);
    const int vr0 = tile_vr0(g, ty), nr = tile_vr1(g, ty) - vr0;
    const int rbase = flip_row(g, vr0);  // memory row of tile row 0; tile row rr is memory row rbase + rr
    const double a_pos = st->a_pos, a_neg = st->a_neg;
    auto classify = [&](double v) -> uint32_t { return v > a_pos ? 1u : ((find_neg && v < a_neg) ? 2u : 0u); };
    auto pix = [&](int cv, int vr) -> double { return ldv(x, size_t(mem_col(g, cv)) * size_t(g.R) + size_t(flip_row(g, vr))); };
    auto mem_index = [&](int c, int rr) -> uint32_t { return uint32_t(mem_col(g, cv0 + c)) * uint32_t(g.R) + uint32_t(rbase + rr); };
    const unsigned long long lt = (1ull << lane) - 1ull;

    // 1. classification and vertical runs (a wave holds one tile column per step)
    uint32_t info[CT_PPT];  // class | run start << 2 | run end << 8
#pragma unroll
    for (int j = 0; j < CT_PPT; j++) {
        const int c = w + CT_WAVES * j, li = c * CT_H + lane;
        uint32_t cl = 0;
        if (c < nc && lane < nr)
            cl = classify(ldv(x, size_t(mem_index(c, lane))));
        const unsigned long long m1 = __ballot(cl == 1u), m2 = __ballot(cl == 2u);
        const unsigned long long M = cl == 2u ? m2 : m1;
        const unsigned long long zb = ~M & lt, za = ~M & ~(lt | (1ull << lane));
        const uint32_t rs = zb ? uint32_t(64 - __builtin_clzll(zb)) : 0u;
        const uint32_t re = za ? uint32_t(__builtin_ctzll(za) - 1) : 63u;
        info[j] = cl | (rs << 2) | (re << 8);
        s_cls[li] = uint8_t(cl);
        s_par[li] = uint32_t(c * CT_H) + (cl ? rs : uint32_t(lane));
        if (lane == 0)
            s_mask[c] = m1 | m2;
    }
    if (t < CT_W) {
        s_rmp[t] = 0;
        s_rmn[t] = 0;
    }
    if (t == 0)
        s_nroot = 0;
    // classes of the frame around the tile (0 outside the image)
    for (int q = t; q < 2 * (CT_H + 2) + 2 * CT_W; q += CT_THREADS) {
        uint32_t cl = 0;
        if (q < 2 * (CT_H + 2)) {
            const bool right = q >= CT_H + 2;
            const int fr = (right ? q - (CT_H + 2) : q) - 1;  // -1 .. CT_H
            const int cv = right ? cv0 + nc : cv0 - 1, vr = vr0 + fr;
            if (fr <= nr && cv >= 0 && cv < g.C && vr >= 0 && vr < g.R)
                cl = classify(pix(cv, vr));
            (right ? s_fr : s_fl)[fr + 1] = uint8_t(cl);
        } else {
            const bool bottom = q >= 2 * (CT_H + 2) + CT_W;
            const int c = q - 2 * (CT_H + 2) - (bottom ? CT_W : 0);
            const int vr = bottom ? vr0 + nr : vr0 - 1;
            if (c < nc && vr >= 0 && vr < g.R)
                cl = classify(pix(cv0 + c, vr));
            (bottom ? s_fb : s_ft)[c] = uint8_t(cl);
        }
    }
    __syncthreads();
    // list of the foreground pixels (key order)
    if (w == 0) {
        const uint32_t cnt = uint32_t(__popcll(s_mask[lane]));
        uint32_t incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off);
            if (lane >= off)
                incl += o;
        }
        s_pre[lane] = incl - cnt;
        if (lane == 63)
            s_pre[64] = incl;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CT_PPT; j++) {
        const int c = w + CT_WAVES * j;
        if (info[j] & 3u)
            s_list[s_pre[c] + uint32_t(__popcll(s_mask[c] & lt))] =
                uint32_t(c * CT_H + lane) | ((info[j] & 3u) << 12) | (((info[j] >> 2) & 63u) << 14) | (((info[j] >> 8) & 63u) << 20);
    }
    const int nfg = int(s_pre[64]);
    __syncthreads();

    // 2. one union per pair of touching runs of neighbouring columns: a run with
    // rows [s, e] meets the left column's rows [s - 1, e + 1] (8-connectivity) or
    // [s, e]; each maximal same-class stretch there is linked once, at its first row
    for (int e = t; e < nfg; e += CT_THREADS) {
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu), c = li >> 6, rr = li & 63;
        const uint32_t cl = (en >> 12) & 3u;
        const int rs = int((en >> 14) & 63u), re = int((en >> 20) & 63u);
        if (c == 0)
            continue;
        const uint32_t a = uint32_t(c * CT_H + rs), lb = uint32_t((c - 1) * CT_H);
        auto lsame = [&](int r) { return r >= 0 && r < nr && s_cls[lb + r] == cl; };
        if (conn4) {
            if (lsame(rr) && (rr == rs || !lsame(rr - 1)))
                lds_union(s_par, a, lb + rr);
        } else {
            if (lsame(rr) && !lsame(rr - 1))
                lds_union(s_par, a, lb + rr);
            if (rr == rs && lsame(rs - 1))
                lds_union(s_par, a, lb + rs - 1);
            if (rr == re && lsame(re + 1) && !lsame(re))
                lds_union(s_par, a, lb + re + 1);
        }
    }
    __syncthreads();
    // 3. roots of the run starts (a component's root, its smallest index, starts a run)
    for (int e = t; e < nfg; e += CT_THREADS) {
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu);
        if ((li & 63) == int((en >> 14) & 63u))
            s_par[li] = lds_find(s_par, li);
    }
    __syncthreads();
    // every pixel points at its root, roots take record slots; components with a
    // same-class neighbour across the tile border are open
    for (int e = t; e < nfg; e += CT_THREADS) {
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu), c = li >> 6, rr = li & 63;
        const uint32_t cl = (en >> 12) & 3u;
        const int rs = int((en >> 14) & 63u);
        const uint32_t root = s_par[c * CT_H + rs] & L_KEY;
        if (rr != rs)
            s_par[li] = root;
        else if (root == uint32_t(li))  // a root: its slot (any order serves)
            atomicOr(&s_par[li], atomicAdd(&s_nroot, 1u) << 16);
        if (!(c == 0 || c == nc - 1 || rr == 0 || rr == nr - 1))
            continue;
        bool open = false;
        if (c == 0)
            open = open || s_fl[rr + 1] == cl || (!conn4 && (s_fl[rr] == cl || s_fl[rr + 2] == cl));
        if (c == nc - 1)
            open = open || s_fr[rr + 1] == cl || (!conn4 && (s_fr[rr] == cl || s_fr[rr + 2] == cl));
        if (rr == 0)
            open = open || s_ft[c] == cl || (!conn4 && ((c > 0 && s_ft[c - 1] == cl) || (c + 1 < nc && s_ft[c + 1] == cl)));
        if (rr == nr - 1)
            open = open || s_fb[c] == cl || (!conn4 && ((c > 0 && s_fb[c - 1] == cl) || (c + 1 < nc && s_fb[c + 1] == cl)));
        if (open)
            atomicOr(&s_par[root], L_OPEN);
    }
    __syncthreads();

    // 4. the closed roots of every segment, the LDS sums of the roots' slots, and the
    // tile's record count (a pool share when it exceeds its region)
    const uint32_t total = s_nroot;
    for (int e = t; e < nfg; e += CT_THREADS) {
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu);
        const uint32_t pv = s_par[li];
        if ((pv & L_KEY) != uint32_t(li))
            continue;
        const uint32_t cl = (en >> 12) & 3u, slot = (pv >> 16) & 0xFFFu;
        if (!(pv & L_OPEN))
            atomicOr(cl == 1u ? &s_rmp[li >> 6] : &s_rmn[li >> 6], 1ull << (li & 63));
        if (slot < uint32_t(CT_RCAP)) {
            r_ext[slot] = cl == 1u ? 0ull : ~0ull;
            r_cnt[slot] = 0;
            r_idx[slot] = ~0u;
#pragma unroll
            for (int q = 0; q < 6; q++)
                r_mom[q][slot] = 0.0;
        }
    }
    if (t == 0) {
        rec.tcount[blockIdx.x] = total;
        s_base = total > rec.rt ? atomicAdd(rec.pool_count, total - rec.rt) : 0u;
    }
    __syncthreads();
    if (w == 1 && lane < nc) {
        const unsigned seg = seg_index(g, cv0 + lane, ty);
        rmask[seg] = s_rmp[lane];
        rmask[g.nseg + seg] = s_rmn[lane];
        rmask[2 * size_t(g.nseg) + seg] = s_mask[lane];  // foreground of either sign (for the label map)
    }
    const uint32_t pool_base = s_base;
    const bool ok = total <= rec.rt || size_t(pool_base) + size_t(total - rec.rt) <= size_t(rec.pool_cap);
    auto ridx = [&](uint32_t slot) -> uint32_t {
        return slot < rec.rt ? blockIdx.x * rec.rt + slot : rec.ntiles * rec.rt + pool_base + (slot - rec.rt);
    };
    if (ok) {
        for (int e = t; e < nfg; e += CT_THREADS) {
            const uint32_t en = s_list[e];
            const int li = int(en & 0xFFFu), c = li >> 6, rr = li & 63;
            const uint32_t pv = s_par[li];
            if ((pv & L_KEY) != uint32_t(li))
                continue;
            const uint32_t cl = (en >> 12) & 3u, slot = (pv >> 16) & 0xFFFu;
            const uint32_t kk = ridx(slot);
            const uint32_t gkey = uint32_t(cv0 + c) * uint32_t(g.R) + uint32_t(rbase + rr);
            rec.key[kk] = gkey | (cl == 2u ? P_NEG : 0u) | ((pv & L_OPEN) ? P_OPEN : 0u);
            rec.pos[kk] = ((seg_index(g, cv0 + c, ty) + (cl == 2u ? g.nseg : 0u)) << 6) | uint32_t(rr);
            if (slot >= uint32_t(CT_RCAP)) {
                IslandSum z{};
                z.ext = cl == 1u ? 0ull : ~0ull;
                z.idx = ~0u;
                rec.sum[kk] = z;
            }
        }
    }
    // the barrier's workgroup-scope release orders these stores before the atomics
    // below: every access of the block meets in the L2 of its XCD
    __syncthreads();

    // 5. island sums: count, extremum, moments (sourcefind.cpp:337-628)
    double ev[CT_PPT];
#pragma unroll
    for (int k = 0; k < CT_PPT; k++) {
        const int e = k * CT_THREADS + t;
        ev[k] = 0.0;
        if (e >= nfg)
            continue;
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu), c = li >> 6, rr = li & 63;
        const uint32_t cl = (en >> 12) & 3u;
        const uint32_t slot = (s_par[s_par[li] & L_KEY] >> 16) & 0xFFFu;
        const double val = ldv(x, size_t(mem_index(c, rr)));
        ev[k] = val;
        const unsigned long long kv = okey(val);
        const double x_idx = double(cv0 + c), y_idx = double(vr0 + rr);
        const double xb = x_idx * val, yb = y_idx * val;
        const double xx = xb * x_idx, yy = yb * y_idx, xy = xb * y_idx;
        if (slot < uint32_t(CT_RCAP)) {
            atomicAdd(&r_cnt[slot], 1);
            if (cl == 1u)
                atomicMax(&r_ext[slot], kv);
            else
                atomicMin(&r_ext[slot], kv);
            unsafeAtomicAdd(&r_mom[0][slot], xb);
            unsafeAtomicAdd(&r_mom[1][slot], yb);
            unsafeAtomicAdd(&r_mom[2][slot], xx);
            unsafeAtomicAdd(&r_mom[3][slot], yy);
            unsafeAtomicAdd(&r_mom[4][slot], xy);
            unsafeAtomicAdd(&r_mom[5][slot], val);
        } else if (ok) {
            IslandSum& s = rec.sum[ridx(slot)];
            atomicAdd(&s.cnt, 1);
            if (cl == 1u)
                atomicMax(&s.ext, kv);
            else
                atomicMin(&s.ext, kv);
            unsafeAtomicAdd(&s.mom[0], xb);
            unsafeAtomicAdd(&s.mom[1], yb);
            unsafeAtomicAdd(&s.mom[2], xx);
            unsafeAtomicAdd(&s.mom[3], yy);
            unsafeAtomicAdd(&s.mom[4], xy);
            unsafeAtomicAdd(&s.mom[5], val);
        }
    }
    __syncthreads();
    // smallest memory index holding the extremum
#pragma unroll
    for (int k = 0; k < CT_PPT; k++) {
        const int e = k * CT_THREADS + t;
        if (e >= nfg)
            continue;
        const uint32_t en = s_list[e];
        const int li = int(en & 0xFFFu), c = li >> 6, rr = li & 63;
        const uint32_t slot = (s_par[s_par[li] & L_KEY] >> 16) & 0xFFFu;
        const unsigned long long kv = okey(ev[k]);
        if (slot < uint32_t(CT_RCAP)) {
            if (kv == r_ext[slot])
                atomicMin(&r_idx[slot], mem_index(c, rr));
        } else if (ok) {
            IslandSum& s = rec.sum[ridx(slot)];
            if (kv == __hip_atomic_load(&s.ext, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(&s.idx, mem_index(c, rr));
        }
    }
    __syncthreads();

    // 6. records of the LDS-summed components; parents of the tile's pixels
    if (ok) {
        for (uint32_t s = t; s < min(total, uint32_t(CT_RCAP)); s += CT_THREADS) {
            IslandSum o;
            o.ext = r_ext[s];
            o.idx = r_idx[s];
            o.cnt = r_cnt[s];
#pragma unroll
            for (int q = 0; q < 6; q++)
                o.mom[q] = r_mom[q][s];
            rec.sum[ridx(s)] = o;
        }
    }
#pragma unroll
    for (int j = 0; j < CT_PPT; j++) {
        const int c = w + CT_WAVES * j;
        if (c >= nc || lane >= nr)
            continue;
        const uint32_t cl = info[j] & 3u;
        const uint32_t gk = uint32_t(cv0 + c) * uint32_t(g.R) + uint32_t(rbase + lane);
        uint32_t out = P_BG;
        if (cl) {
            const uint32_t root = s_par[c * CT_H + lane] & L_KEY;
            const uint32_t rk = uint32_t(cv0 + int(root / CT_H)) * uint32_t(g.R) + uint32_t(rbase + int(root % CT_H));
            out = rk | (cl == 2u ? P_NEG : 0u) | ((s_par[root] & L_OPEN) ? P_OPEN : 0u);
        }
        P[gk] = out;
    }
}

// Unions across the tile borders: the tile's first column with the column on its
// left (threads 0..63) and its first row with the row above (threads 64..127).
__global__ __launch_bounds__(128) void ccl_border(uint32_t* P, TileGeom g, int conn4)
{
    const int t = threadIdx.x;
    const int tx = int(blockIdx.x) % g.ntx, ty = int(blockIdx.x) / g.ntx;
    const int cv0 = tx * CT_W, nc = min(CT_W, g.C - cv0);
    const int vr0 = tile_vr0(g, ty), nr = tile_vr1(g, ty) - vr0;
    int cv, vr;
    if (t < 64) {
        if (t >= nr || cv0 == 0)
            return;
        cv = cv0;
        vr = vr0 + t;
    } else {
        if (t - 64 >= nc || vr0 == 0)
            return;
        cv = cv0 + (t - 64);
        vr = vr0;
    }
    const uint32_t a = uint32_t(cv) * uint32_t(g.R) + uint32_t(flip_row(g, vr));
    const uint32_t pa = P[a];
    if (pa == P_BG)
        return;
    const uint32_t flags = pa & (P_OPEN | P_NEG);
    for (int d = -1; d <= 1; d++) {
        if (conn4 && d != 0)
            continue;
        int ncv, nvr;
        if (t < 64) {
            ncv = cv0 - 1;
            nvr = vr + d;
        } else {
            ncv = cv + d;
            nvr = vr0 - 1;
        }
        if (ncv < 0 || ncv >= g.C || nvr < 0 || nvr >= g.R)
            continue;
        const uint32_t b = uint32_t(ncv) * uint32_t(g.R) + uint32_t(flip_row(g, nvr));
        const uint32_t pb = P[b];
        if (pb == P_BG || (pb & P_NEG) != (pa & P_NEG))
            continue;
        g_union(P, a, b, flags);
    }
}

// Open components: local root -> global root; global roots join the root masks.
__global__ __launch_bounds__(256) void ccl_open_roots(uint32_t* P, TileGeom g, CompRecords rec, unsigned long long* rmask)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (pool_overflow(rec) || !rec_valid(rec, i))
        return;
    const uint32_t k = rec.key[i];
    if (!(k & P_OPEN))
        return;
    const uint32_t lr = k & P_KEY;
    const uint32_t gr = g_find(P, lr);
    P[lr] = gr | (k & (P_OPEN | P_NEG));
    if (gr == lr) {
        int rr;
        const unsigned sseg = seg_of_key(g, lr, k & P_NEG, rr);
        atomicOr(&rmask[sseg], 1ull << rr);
    }
}

// Label counts, and whether the island stages can run: every tile component has
// a record (so the labels fit the cap of the island arrays) and the median held.
struct LabelCounts {
    SFState* st;
    const SelState* ss;
    unsigned long long cap;  // labels the island arrays hold
};

// Open components whose local root is the global root start their island's sums;
// thread 0 also sets the label counts (read by the later island kernels).
__global__ __launch_bounds__(256) void ccl_open_init(const uint32_t* P, SegIds s, CompRecords rec, IslandSum* isum,
    uint32_t* lrec, LabelCounts lc)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        SFState* st = lc.st;
        const uint32_t nrec = *rec.pool_count, kp = s.offs[s.g.nseg], kall = s.offs[2 * s.g.nseg];
        st->num_pos = kp;
        st->num_neg = kall - kp;
        st->num_records = nrec;
        st->fallback = lc.ss->fallback;
        st->skip = nrec > rec.pool_cap || lc.ss->fallback || (unsigned long long)kall > lc.cap;
    }
    if (pool_overflow(rec) || !rec_valid(rec, i))
        return;
    const uint32_t k = rec.key[i];
    const uint32_t lr = k & P_KEY;
    if (!(k & P_OPEN) || (P[lr] & P_KEY) != lr)
        return;
    const bool neg = k & P_NEG;
    IslandSum z{};
    z.ext = neg ? ~0ull : 0ull;
    z.idx = ~0u;
    const uint32_t gid = root_gid(s, lr, neg);
    isum[gid] = z;
    lrec[gid] = LREC_OPEN;
}

// Island sums by list position: a closed component's label points at its record,
// open ones are combined into the island's sums.
__global__ __launch_bounds__(256) void ccl_records(const uint32_t* P, SegIds s, CompRecords rec, IslandSum* isum,
    uint32_t* lrec)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (pool_overflow(rec) || !rec_valid(rec, i))
        return;
    const uint32_t k = rec.key[i];
    if (!(k & P_OPEN)) {
        const uint32_t pos = rec.pos[i];
        lrec[gid_at(s, pos >> 6, int(pos & 63u))] = i;
        return;
    }
    const bool neg = k & P_NEG;
    IslandSum& d = isum[root_gid(s, P[k & P_KEY] & P_KEY, neg)];
    const IslandSum& r = rec.sum[i];
    if (neg)
        atomicMin(&d.ext, r.ext);
    else
        atomicMax(&d.ext, r.ext);
    atomicAdd(&d.cnt, r.cnt);
    for (int q = 0; q < 6; q++)
        unsafeAtomicAdd(&d.mom[q], r.mom[q]);
}

// Open islands: smallest memory index among the components holding the extremum.
__global__ __launch_bounds__(256) void ccl_open_idx(const uint32_t* P, SegIds s, CompRecords rec, IslandSum* isum)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (pool_overflow(rec) || !rec_valid(rec, i))
        return;
    const uint32_t k = rec.key[i];
    if (!(k & P_OPEN))
        return;
    IslandSum& d = isum[root_gid(s, P[k & P_KEY] & P_KEY, k & P_NEG)];
    if (rec.sum[i].ext == d.ext)
        atomicMin(&d.idx, rec.sum[i].idx);
}

constexpr int LM_THREADS = 256, LM_WAVES = LM_THREADS / 64, LM_PPT = CT_PIX / LM_THREADS;

// Label map in memory order (LM_THREADS threads per tile): +-(id + 1), 0 for background and, with gen_map,
// for islands whose extremum misses the detection threshold (sourcefind.cpp:462-492).
__global__ __launch_bounds__(LM_THREADS) void ccl_label_map(const uint32_t* __restrict__ P, SegIds s, const SFState* st,
    const IslandSum* __restrict__ isum, const uint32_t* __restrict__ lrec, const IslandSum* __restrict__ rsum, int gen_map,
    int32_t* __restrict__ lab)
{
    const TileGeom& g = s.g;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tx = int(blockIdx.x) % g.ntx, ty = int(blockIdx.x) / g.ntx;
    const int cv0 = tx * CT_W, nc = min(CT_W, g.C - cv0);
    const int vr0 = tile_vr0(g, ty), nr = tile_vr1(g, ty) - vr0;
    const int rbase = flip_row(g, vr0);
    if (lane >= nr || st->skip)
        return;
    const double d_pos = st->d_pos, d_neg = st->d_neg;
    const uint32_t kp = s.offs[g.nseg];
    // Each phase only issues loads into registers (none waits on another of its
    // phase): parents, roots of open components, segment masks and offsets, extrema.
    // a pixel's parent is read only where the tile labelling found foreground (the
    // segment's mask, one uniform load per column): background needs no load
    uint32_t pv[LM_PPT], rv[LM_PPT];
    unsigned long long fm[LM_PPT];
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        const int c = w + LM_WAVES * j;
        fm[j] = c < nc ? s.mask[2 * size_t(g.nseg) + seg_index(g, cv0 + c, ty)] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        const int c = w + LM_WAVES * j;
        pv[j] = ((fm[j] >> lane) & 1ull) ? P[uint32_t(cv0 + c) * uint32_t(g.R) + uint32_t(rbase + lane)] : P_BG;
    }
#pragma unroll
    for (int j = 0; j < LM_PPT; j++)
        rv[j] = (pv[j] != P_BG && (pv[j] & P_OPEN)) ? P[pv[j] & P_KEY] : pv[j];
    uint32_t so[LM_PPT];
    unsigned long long sm[LM_PPT];
    int rr[LM_PPT];
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        const unsigned sseg = pv[j] != P_BG ? seg_of_key(g, rv[j] & P_KEY, pv[j] & P_NEG, rr[j]) : 0u;
        so[j] = pv[j] != P_BG ? s.offs[sseg] : 0u;
        sm[j] = pv[j] != P_BG ? s.mask[sseg] : 0ull;
    }
    uint32_t lr[LM_PPT];
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        so[j] += uint32_t(__popcll(sm[j] & ((1ull << rr[j]) - 1ull)));  // the island id
        lr[j] = (gen_map && pv[j] != P_BG) ? lrec[so[j]] : 0u;
    }
    unsigned long long ev[LM_PPT];
#pragma unroll
    for (int j = 0; j < LM_PPT; j++)
        ev[j] = (gen_map && pv[j] != P_BG) ? (lr[j] == LREC_OPEN ? isum[so[j]].ext : rsum[lr[j]].ext) : 0ull;
    // the labels, then the stores (a wait for a load between stores would also
    // wait for the stores before it)
    int32_t L[LM_PPT];
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        L[j] = 0;
        if (pv[j] != P_BG) {
            const bool neg = pv[j] & P_NEG;
            const uint32_t gid = so[j];
            bool keep = true;
            if (gen_map) {
                const double e = okey_inv(ev[j]);
                keep = neg ? (e < d_neg) : (e > d_pos);
            }
            L[j] = keep ? (neg ? -int32_t(gid - kp + 1) : int32_t(gid + 1)) : 0;
        }
    }
#pragma unroll
    for (int j = 0; j < LM_PPT; j++) {
        const int c = w + LM_WAVES * j;
        if (c < nc)
            lab[size_t(mem_col(g, cv0 + c)) * size_t(g.R) + size_t(rbase + lane)] = L[j];
    }
}

// ------------------------------------------------------------------ islands
// IslandParams::estimate_moments_fit (sourcefind.cpp:691-717) and
// Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57).
__device__ void moments_fit(stp_island& isl, double x_bar, double y_bar, double xx_bar, double yy_bar, double xy_bar,
    double rms_est, double analysis_n_sigma)
{
    const double working1 = (xx_bar + yy_bar) / 2.0;
    const double working2 = sqrt(((xx_bar - yy_bar) * (xx_bar - yy_bar) / 4) + xy_bar * xy_bar);
    const double major_sq = working1 + working2, minor_sq = working1 - working2;
    const double pixel_threshold = analysis_n_sigma * rms_est;
    const double cutoff_ratio = isl.sign * isl.extremum_val / pixel_threshold;
    const double axes_scale = 1.0 - log(cutoff_ratio) / (cutoff_ratio - 1.0);
    double semimajor = sqrt(major_sq / axes_scale), semiminor = sqrt(minor_sq / axes_scale);
    double theta = 0.5 * atan(2.0 * xy_bar / (xx_bar - yy_bar));
    if (theta * xy_bar < 0.0)
        theta += M_PI / 2.0;
    const double half_pi = M_PI / 2.0;
    semimajor = fabs(semimajor);
    semiminor = fabs(semiminor);
    if (semimajor < semiminor) {
        const double t = semimajor;
        semimajor = semiminor;
        semiminor = t;
        theta += half_pi;
    }
    double mt = fmod(theta, M_PI);
    if (mt <= -half_pi)
        mt += M_PI;
    else if (mt > half_pi)
        mt -= M_PI;
    isl.moments_fit.amplitude = isl.extremum_val;
    isl.moments_fit.x_centre = x_bar;
    isl.moments_fit.y_centre = y_bar;
    isl.moments_fit.semimajor = semimajor;
    isl.moments_fit.semiminor = semiminor;
    isl.moments_fit.theta = mt;
}

// Label g of the concatenated list: positive labels 1..kp first, then negative
// labels -1..-kn (the island order of sourcefind.cpp:235-278).
// valid: extremum beyond the detection threshold (sourcefind.cpp:417-436);
// keep: valid and at least min_area samples (sourcefind.cpp:240,263).
// The grid covers cap >= kp + kn labels (kp, kn from the state); keep is 0 beyond.
// vcnt / kcnt: per block of 256 labels, the valid and the kept labels.
__global__ __launch_bounds__(256) void sf_island_flags(const IslandSum* isum, const uint32_t* lrec, const IslandSum* rsum,
    uint32_t cap, const SFState* st, int min_area, uint32_t* keep, uint32_t* vcnt, uint32_t* kcnt)
{
    __shared__ uint32_t wv[4], wk[4];
    if (st->skip)
        return;
    const uint32_t kp = st->num_pos, kn = st->num_neg;
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    bool v = false, k = false;
    if (g < kp + kn) {
        const bool pos = g < kp;
        const IslandSum& sm = label_sum(isum, lrec, rsum, g);
        const double ext = okey_inv(sm.ext);
        v = pos ? (ext > st->d_pos) : (ext < st->d_neg);
        k = v && sm.cnt >= min_area;
        keep[g] = k ? 1u : 0u;
    } else if (g < cap) {
        keep[g] = 0u;
    }
    const uint32_t c = uint32_t(__popcll(__ballot(v))), ck = uint32_t(__popcll(__ballot(k)));
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = c;
        wk[threadIdx.x >> 6] = ck;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        vcnt[blockIdx.x] = wv[0] + wv[1] + wv[2] + wv[3];
        kcnt[blockIdx.x] = wk[0] + wk[1] + wk[2] + wk[3];
    }
}

// One island record per kept label at its rank among the kept labels; block 0
// also writes the island count and the count of valid labels (vcnt: one per block
// of sf_island_flags, the same grid).
// One island record (sourcefind.cpp:235-278, 592-613) from the sums of label g.
__device__ void emit_island(const IslandSum& s, uint32_t g, uint32_t kp, int rows, int cols, double rms,
    double analysis_n_sigma, stp_island* dst)
{
    const bool pos = g < kp;
    const uint32_t i = pos ? g : g - kp;
    const size_t li = size_t(s.idx);
    const int r = int(li % size_t(rows)), c = int(li / size_t(rows));
    const int h_shift = cols / 2, v_shift = rows / 2;
    stp_island isl{};
    isl.label_idx = pos ? int32_t(i + 1) : -int32_t(i + 1);
    isl.sign = pos ? 1 : -1;
    isl.extremum_y_idx = r < v_shift ? r + v_shift : r - v_shift;
    isl.extremum_x_idx = c < h_shift ? c + h_shift : c - h_shift;
    isl.num_samples = s.cnt;
    isl.extremum_val = okey_inv(s.ext);
    // moments normalised by the flux sum (sourcefind.cpp:594-613)
    const double* m = s.mom;
    const double xb = m[0] / m[5], yb = m[1] / m[5];
    moments_fit(isl, xb, yb, m[2] / m[5] - xb * xb, m[3] / m[5] - yb * yb, m[4] / m[5] - xb * yb, rms,
        analysis_n_sigma);
    *dst = isl;
}

// out holds out_cap records (device memory, or page-locked host memory written over
// the bus); islands beyond it are counted, not written.  A block's islands form one
// run of the list: they are assembled in LDS and written as consecutive 16-byte
// chunks by consecutive lanes (whole lines on the bus).
static_assert(sizeof(stp_island) % 16 == 0, "island records of whole 16-byte chunks");
constexpr int ISL_CHUNKS = int(sizeof(stp_island) / 16);
// A block's islands start after the kept labels of the blocks before it (kcnt, added
// up by the block itself) and follow in label order (wave ballots of keep).
// Block 0 also writes the final state to host_st (page-locked host memory, over the
// bus) when it is given, also when the labelling is incomplete: the state the host
// reads after the call, with no copy launch.
constexpr unsigned SF_STATE_WORDS = unsigned(sizeof(SFState) / 4);
static_assert(sizeof(SFState) % 4 == 0, "state of whole 32-bit words");
__global__ __launch_bounds__(256) void sf_island_emit(const IslandSum* isum, const uint32_t* lrec, const IslandSum* rsum,
    const uint32_t* keep, const uint32_t* kcnt, const uint32_t* vcnt, SFState* st, int rows, int cols,
    double analysis_n_sigma, stp_island* out, uint32_t out_cap, SFState* host_st)
{
    __shared__ uint32_t red[256], red2[256];
    __shared__ uint32_t wk[4];
    __shared__ uint4 s_isl[256 * ISL_CHUNKS];
    const int skip = st->skip;
    if (skip) {
        if (blockIdx.x == 0 && host_st)
            for (unsigned t = threadIdx.x; t < SF_STATE_WORDS; t += 256)
                ((uint32_t*)host_st)[t] = ((const uint32_t*)st)[t];
        return;
    }
    const uint32_t kp = st->num_pos, kn = st->num_neg;
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0) {
        // the blocks holding labels (unconditional loads with a clamped index, so
        // that a batch is in flight together)
        const uint32_t nb = (kp + kn + 255) / 256;
        uint32_t c = 0, ck = 0;
        for (uint32_t b0 = 0; b0 < nb; b0 += 256 * 8) {
            uint32_t v[8], u[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                v[j] = vcnt[min(b0 + 256 * j + threadIdx.x, nb > 0 ? nb - 1 : 0u)];
                u[j] = kcnt[min(b0 + 256 * j + threadIdx.x, nb > 0 ? nb - 1 : 0u)];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                c += b0 + 256 * j + threadIdx.x < nb ? v[j] : 0u;
                ck += b0 + 256 * j + threadIdx.x < nb ? u[j] : 0u;
            }
        }
        red[threadIdx.x] = c;
        red2[threadIdx.x] = ck;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (threadIdx.x < off) {
                red[threadIdx.x] += red[threadIdx.x + off];
                red2[threadIdx.x] += red2[threadIdx.x + off];
            }
            __syncthreads();
        }
        const uint32_t nv = red[0], ni = red2[0];
        if (threadIdx.x == 0) {
            st->num_valid = nv;
            st->num_islands = ni;
        }
        if (host_st) {
            // the two counts come from this block's sums, every other word is final
            constexpr unsigned WV = unsigned(offsetof(SFState, num_valid) / 4);
            constexpr unsigned WI = unsigned(offsetof(SFState, num_islands) / 4);
            for (unsigned t = threadIdx.x; t < SF_STATE_WORDS; t += 256) {
                const uint32_t v = ((const uint32_t*)st)[t];
                ((uint32_t*)host_st)[t] = t == WV ? nv : t == WI ? ni : v;
            }
        }
        __syncthreads();  // red is reused below
    }
    const uint32_t nlab = kp + kn, g0 = blockIdx.x * 256;
    if (g0 >= nlab)
        return;
    // kept islands of the blocks before this one (loads in flight together)
    uint32_t pb = 0;
    for (uint32_t b0 = 0; b0 < blockIdx.x; b0 += 256 * 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = kcnt[min(b0 + 256 * j + threadIdx.x, blockIdx.x - 1)];
#pragma unroll
        for (int j = 0; j < 8; j++)
            pb += b0 + 256 * j + threadIdx.x < blockIdx.x ? v[j] : 0u;
    }
    red[threadIdx.x] = pb;
    // the position of this label's island within the block
    const bool kept = g < nlab && keep[g];
    const unsigned long long bal = __ballot(kept);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        wk[w] = uint32_t(__popcll(bal));
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const uint32_t base = red[0], cnt = wk[0] + wk[1] + wk[2] + wk[3];
    uint32_t pos = uint32_t(__popcll(bal & ((1ull << lane) - 1ull)));
    for (int k = 0; k < w; k++)
        pos += wk[k];
    if (kept)
        emit_island(label_sum(isum, lrec, rsum, g), g, kp, rows, cols, st->rms, analysis_n_sigma,
            (stp_island*)s_isl + pos);
    __syncthreads();
    uint4* o = (uint4*)out + size_t(base) * ISL_CHUNKS;
    for (uint32_t q = threadIdx.x; q < cnt * ISL_CHUNKS; q += 256)
        if (base + q / ISL_CHUNKS < out_cap)
            o[q] = s_isl[q];
}


__global__ void sf_fallback_clear(SelState* ss)
{
    ss->fallback = 0;
}

// ------------------------------------------------------------------ host helpers
template <typename U>
U* wsp(stp_ctx* ctx, const char* name, size_t count)
{
    return (U*)ws_get(ctx, name, sizeof(U) * std::max<size_t>(count, 1));
}

void d2h_sync(stp_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// streaming passes: HWG threads per block, one block per 64 Ki elements, at most 256 blocks
unsigned hist_blocks(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>(256, (n + 65535) / 65536))); }


__global__ void sf_write_value(double* out, unsigned long long key)
{
    out[0] = okey_inv(key);
}

// Exact order statistic x_(k) (0-based) by radix selection on the order keys,
// 8 bits at a time, until the candidates fit one workgroup sort.  Writes d_out[0].
template <typename T>
void radix_select(stp_ctx* ctx, const T* x, size_t n, size_t k, double* d_out)
{
    const unsigned nb = hist_blocks(n);
    uint32_t* rbh = wsp<uint32_t>(ctx, "sf_rbh", size_t(nb) * 256);
    unsigned long long* rhist = wsp<unsigned long long>(ctx, "sf_rhist", 256);
    unsigned long long* cand = wsp<unsigned long long>(ctx, "sf_cand", SORT_CAP);
    unsigned* counter = wsp<unsigned>(ctx, "sf_counter", 1);
    double* tmp = wsp<double>(ctx, "sf_sel_tmp", 2);
    std::vector<unsigned long long> rh(256);
    unsigned long long prefix = 0;
    int shift = 64;
    size_t before = 0, cnt = n;
    while (cnt > size_t(SORT_CAP) && shift > 0) {
        shift -= 8;
        launch(ctx, "sf_radix_hist", sf_radix_hist<T>, dim3(nb), dim3(HWG), 0, x, n, shift, prefix, int(shift < 56), rbh);
        fill_u32(ctx, (uint32_t*)rhist, 0u, 2 * 256);
        launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(1, std::min(64u, nb)), dim3(256), 0, (const uint32_t*)rbh, int(nb),
            256, rhist, (const int*)nullptr);
        d2h_sync(ctx, rh.data(), rhist, sizeof(unsigned long long) * 256);
        size_t cum = before;
        int d = 0;
        for (; d < 256; d++) {
            if (cum + rh[d] > k)
                break;
            cum += rh[d];
        }
        if (d == 256)
            throw Error(STP_ERR_INTERNAL, "radix selection lost the target rank");
        prefix = (prefix << 8) | unsigned(d);
        before = cum;
        cnt = rh[d];
    }
    if (cnt > size_t(SORT_CAP)) {
        // every candidate has the same 64-bit key: the value is known exactly
        launch(ctx, "sf_select", sf_write_value, dim3(1), dim3(1), 0, d_out, prefix);
        return;
    }
    GatherSpec gs{};
    gs.shift = shift;
    gs.prefix = prefix;
    fill_u32(ctx, counter, 0u, 1);
    launch(ctx, "sf_gather", sf_gather<T>, dim3(nb), dim3(HWG), 0, x, n, gs, cand, counter, unsigned(SORT_CAP));
    launch(ctx, "sf_sort_select", sf_sort_select, dim3(1), dim3(1024), 0, (const unsigned long long*)cand,
        (const unsigned*)counter, int(k - before), int(k - before), tmp);
    HIP_CHECK(hipMemcpyAsync(d_out, tmp, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

// Enqueues the device-driven exact median search (SelState).  It runs when
// st->need_exact is set on the device (force sets it), reads st->mean and
// st->sigma, and sets st->median; ss->fallback reports data the bins cannot
// resolve.  No host round trip.
// Selection set-up of the exact median of n values (buffers of the search).
SelInit sel_setup(stp_ctx* ctx, size_t n, int force)
{
    SelInit si{};
    si.enabled = 1;
    si.k_hi = n / 2;
    si.k_lo = (n & 1) ? n / 2 : n / 2 - 1;
    si.odd = int(n & 1);
    si.force = force;
    si.hist1 = wsp<unsigned long long>(ctx, "sf_hist", LIN_BINS + 2);
    si.hist2 = wsp<unsigned long long>(ctx, "sf_hist2", LIN_BINS);
    si.counter = wsp<unsigned>(ctx, "sf_counter", 1);
    return si;
}

// Enqueues the device-driven exact median search (SelState).  It runs when
// st->need_exact is set on the device (si.force sets it), reads st->mean and
// st->sigma, and sets st->median; ss->fallback reports data the bins cannot
// resolve.  No host round trip.  initialised: the set-up already ran (fused into
// the statistics kernel).
template <typename T>
void enqueue_exact_median(stp_ctx* ctx, const T* x, size_t n, SFState* st, SelState* ss, const SelInit& si, bool initialised)
{
    const unsigned nb = hist_blocks(n);
    uint32_t* bh = wsp<uint32_t>(ctx, "sf_bh", size_t(nb) * (LIN_BINS + 2));
    unsigned long long* cand = wsp<unsigned long long>(ctx, "sf_cand", SORT_CAP);
    const unsigned slices = std::min(16u, nb);
    if (!initialised)
        launch(ctx, "sel_init", sel_init, dim3(1), dim3(256), 0, st, ss, si);
    launch(ctx, "sel_hist1", sel_hist1<T>, dim3(nb), dim3(HWG), 0, x, n, (const SelState*)ss, bh);
    launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(LIN_BINS + 2, 256), slices), dim3(256), 0,
        (const uint32_t*)bh, int(nb), LIN_BINS + 2, si.hist1, (const int*)&ss->run1);
    launch(ctx, "sel_pick1", sel_pick1, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)si.hist1);
    launch(ctx, "sel_hist2", sel_hist2<T>, dim3(nb), dim3(HWG), 0, x, n, (const SelState*)ss, bh);
    launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(LIN_BINS, 256), slices), dim3(256), 0,
        (const uint32_t*)bh, int(nb), LIN_BINS, si.hist2, (const int*)&ss->run2);
    launch(ctx, "sel_pick2", sel_pick2, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)si.hist2);
    launch(ctx, "sel_gather", sel_gather<T>, dim3(nb), dim3(HWG), 0, x, n, (const SelState*)ss, cand, si.counter);
    launch(ctx, "sel_sort", sel_sort, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)cand, (const unsigned*)si.counter,
        st);
}

// Host-driven fallback: radix selection of both ranks, then the median.
template <typename T>
void radix_median(stp_ctx* ctx, const T* x, size_t n, SFState* st, double* sel)
{
    const size_t k_hi = n / 2, k_lo = (n & 1) ? n / 2 : n / 2 - 1;
    radix_select<T>(ctx, x, n, k_lo, sel);
    radix_select<T>(ctx, x, n, k_hi, sel + 1);
    launch(ctx, "sf_median", sf_set_median, dim3(1), dim3(1), 0, st, (const double*)sel, int(n & 1));
}

template <typename T>
void source_find_impl(stp_ctx* ctx, const T* x, int rows, int cols, const stp_sourcefind_pars& p, int32_t* d_lab_out,
    stp_island* islands, size_t cap, size_t* n_islands, stp_sourcefind_stats* stats)
{
    STP_REQUIRE(rows >= 2 && cols >= 2 && rows % 2 == 0 && cols % 2 == 0, "rows and cols must be even (ccl.h:116-117)");
    STP_REQUIRE(rows <= 65536 && cols <= 65536, "rows and cols must be <= 65536 (ccl.h:118-119)");
    const size_t n = size_t(rows) * size_t(cols);
    STP_REQUIRE(n < (size_t(1) << 30), "image too large for 30-bit label keys");
    if (p.gaussian_fitting)
        throw Error(STP_ERR_UNSUPPORTED, "gaussian_fitting: fit the islands with stp_fit_islands[_device] on this "
                                         "call's image and label map (the island record has no fit member)");
    STP_REQUIRE(p.median_method >= STP_MEDIAN_ZEROMEDIAN && p.median_method <= STP_MEDIAN_NTHELEMENT, "unknown median method");

    // device state of the search: SFState then SelState, cleared by one fill
    static_assert(sizeof(SFState) % 8 == 0 && sizeof(SelState) % 8 == 0, "state sizes");
    char* stbuf = wsp<char>(ctx, "sf_state", sizeof(SFState) + sizeof(SelState));
    SFState* st = (SFState*)stbuf;
    SelState* ss = (SelState*)(stbuf + sizeof(SFState));
    double* sel = wsp<double>(ctx, "sf_sel", 2);
    const size_t nblk = (n + CANON_BLOCK - 1) / CANON_BLOCK;
    double* part = wsp<double>(ctx, "sf_part", 2 * nblk + 2);
    double* scr = wsp<double>(ctx, "sf_scr", canon_scratch_doubles(nblk));
    const bool need_rms = !(std::fabs(p.rms_est) > 0.0);
    const int mm = p.median_method;
    bool stats_valid = false;
    // the state starts zeroed: by the first statistics kernel, else by a fill
    const unsigned nstate = unsigned((sizeof(SFState) + sizeof(SelState)) / sizeof(double));
    bool cleared = false;
    auto clear_state = [&]() {
        if (!cleared)
            fill_f64(ctx, (double*)stbuf, 0.0, nstate);
        cleared = true;
    };
    // mean and sigma (canonical sums); si.enabled also sets up the exact-median search
    auto compute_mean_sigma = [&](const SelInit& si) {
        launch(ctx, "sf_sum", sf_sum_kernel<T>, dim3(div_up(nblk, CANON_GROUP)), dim3(SWG), 0, x, n, nblk, part);
        size_t mfin = 0;
        const double* fin = canon_reduce_to_block(ctx, part, nblk, scr, &mfin);
        launch(ctx, "sf_stats", sf_stats, dim3(1), dim3(CANON_LANES), 0, fin, unsigned(mfin), (unsigned long long)n, st, ss, si,
            cleared ? (double*)nullptr : (double*)stbuf, nstate);
        cleared = true;
        stats_valid = true;
    };
    // median (sourcefind.cpp:182-197), enqueued without host round trips: the
    // BINAPPROX pick and the exact search are gated on the device state
    if (mm == STP_MEDIAN_BINAPPROX) {
        compute_mean_sigma(SelInit{});
        launch(ctx, "sf_ba_gate", sf_ba_gate, dim3(1), dim3(1), 0, st);
        const unsigned nb = hist_blocks(n);
        uint32_t* bh = wsp<uint32_t>(ctx, "sf_babh", size_t(nb) * (BA_BINS + 1));
        unsigned long long* hist = wsp<unsigned long long>(ctx, "sf_bahist", BA_BINS + 1);
        launch(ctx, "sf_binapprox_hist", sf_binapprox_hist<T>, dim3(nb), dim3(HWG), 0, x, n, (const SFState*)st, bh);
        fill_u32(ctx, (uint32_t*)hist, 0u, 2 * size_t(BA_BINS + 1));
        launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(BA_BINS + 1, 256), std::min(64u, nb)), dim3(256), 0,
            (const uint32_t*)bh, int(nb), BA_BINS + 1, hist, (const int*)&st->ba_active);
        launch(ctx, "sf_binapprox_pick", sf_binapprox_pick, dim3(1), dim3(1), 0, (const unsigned long long*)hist,
            (unsigned long long)n, st);
        enqueue_exact_median<T>(ctx, x, n, st, ss, sel_setup(ctx, n, 0), false);
    } else if (mm == STP_MEDIAN_BINMEDIAN || mm == STP_MEDIAN_NTHELEMENT) {
        // mean and sigma seed the selection's linear bins (and estimate_rms,
        // sourcefind.cpp:30-36)
        const SelInit si = sel_setup(ctx, n, 1);
        compute_mean_sigma(si);
        enqueue_exact_median<T>(ctx, x, n, st, ss, si, true);
    }
    // ZEROMEDIAN: the state's median stays 0
    if (need_rms && !stats_valid)
        compute_mean_sigma(SelInit{});
    clear_state();  // no statistics kernel ran (ZEROMEDIAN with a given RMS)

    // RMS (sourcefind.cpp:206) and thresholds (:298-301)
    // reset: a counter of the labelling cleared with the thresholds
    auto rms_and_thresholds = [&](uint32_t* reset) {
        const ThreshArgs ta{ int(need_rms), p.rms_est, p.detection_n_sigma, p.analysis_n_sigma, reset };
        const int iters = need_rms ? std::max(0, p.sigma_clip_iters) : 0;
        if (iters > 0) {
            uint32_t* removed = wsp<uint32_t>(ctx, "sf_removed", nblk);
            double* cpart = wsp<double>(ctx, "sf_cpart", 2 * nblk + 2);
            T* cand = wsp<T>(ctx, "sf_clip_cand", nblk * CANON_BLOCK);
            uint32_t* ltab = wsp<uint32_t>(ctx, "sf_clip_lanes", nblk * CANON_LANES);
            for (int it = 0; it < iters; it++) {
                if (it == 0)
                    launch(ctx, "sf_clip", sf_clip_kernel<T>, dim3(div_up(nblk, CANON_GROUP)), dim3(SWG), 0, x, n, nblk,
                        (const SFState*)st, 3.0, cpart, removed, cand, ltab);
                else
                    launch(ctx, "sf_clip_next", sf_clip_next_kernel<T>, dim3(div_up(nblk, SWG / 64)), dim3(SWG), 0, x, n,
                        nblk, (const SFState*)st, cpart, removed, cand, ltab);
                size_t mfin = 0;
                const double* fin = canon_reduce_to_block(ctx, cpart, nblk, scr, &mfin);
                launch(ctx, "sf_clip_update", sf_clip_update, dim3(1), dim3(256), 0, st, fin, unsigned(mfin),
                    (const uint32_t*)removed, int(nblk), it, (unsigned long long)n, iters, 3.0, int(it == iters - 1), ta);
            }
        } else {
            if (need_rms)
                launch(ctx, "sf_clip_init", sf_clip_init, dim3(1), dim3(1), 0, st, (unsigned long long)n, 0, 3.0);
            launch(ctx, "sf_thresholds", sf_thresholds, dim3(1), dim3(1), 0, st, ta);
        }
    };

    // connected components and island sums by tiles (ccl.h:107-773, sourcefind.cpp:337-628)
    const TileGeom tg = tile_geom(rows, cols);
    const unsigned ntiles = unsigned(tg.ntx) * unsigned(tg.nty);
    const size_t nseg2 = 2 * size_t(tg.nseg);
    uint32_t* P = wsp<uint32_t>(ctx, "ccl_parent", n);
    unsigned long long* rmask = wsp<unsigned long long>(ctx, "ccl_rmask", nseg2 + tg.nseg);
    uint32_t* segoff = wsp<uint32_t>(ctx, "ccl_segoff", nseg2 + 1);
    const int find_neg = int(p.find_negative_sources != 0), conn4 = int(p.ccl_4connectivity != 0);
    // component records: a region of CT_RT per tile, then a pool for tiles with more
    uint32_t* tcount = wsp<uint32_t>(ctx, "ccl_tcount", ntiles);
    uint32_t* pool_count = wsp<uint32_t>(ctx, "ccl_npool", 1);
    CompRecords rec{};
    auto records = [&](size_t want_pool) {
        constexpr size_t per = sizeof(IslandSum) + 2 * sizeof(uint32_t);
        const size_t region = size_t(ntiles) * CT_RT;
        const size_t have = ws_bytes(ctx, "ccl_records") / per;
        const size_t pool = std::max(want_pool, have > region ? have - region : size_t(0));
        STP_REQUIRE(region + pool < (size_t(1) << 32), "image too large for 32-bit component records");
        const size_t c = region + pool;
        char* b = (char*)ws_get(ctx, "ccl_records", c * per);
        rec.sum = (IslandSum*)b;
        rec.key = (uint32_t*)(b + sizeof(IslandSum) * c);
        rec.pos = rec.key + c;
        rec.tcount = tcount;
        rec.pool_count = pool_count;
        rec.rt = CT_RT;
        rec.ntiles = ntiles;
        rec.pool_cap = uint32_t(pool);
    };
    records(std::max<size_t>(4096, n / 256));
    // clear_pool: the pool counter was not cleared with the thresholds
    auto label = [&](bool clear_pool) {
        if (clear_pool)
            fill_u32(ctx, pool_count, 0u, 1);
        launch(ctx, "ccl_tile", ccl_tile<T>, dim3(ntiles), dim3(CT_THREADS), 0, x, tg, (const SFState*)st, find_neg, conn4, P,
            rmask, rec);
        launch(ctx, "ccl_border", ccl_border, dim3(ntiles), dim3(128), 0, P, tg, conn4);
        launch(ctx, "ccl_open_roots", ccl_open_roots, dim3(div_up(size_t(ntiles) * CT_RT + rec.pool_cap, 256)), dim3(256), 0, P,
            tg, rec, rmask);
        scan_popc64(ctx, rmask, nseg2, segoff);
    };
    // island sums in list order (positive labels, then negative); label map
    // (sourcefind.cpp:337-628); the island list: validity, area cut, compaction and
    // moments fit.  Sized for one label per record (the label counts stay on the
    // device: no host round trip), skipped by the kernels when the labelling is
    // incomplete.
    stp_island* disl = nullptr;
    // a page-locked caller buffer is written by the emitting kernel itself
    stp_island* islands_dev = (stp_island*)pinned_device_ptr(islands);
    const bool pinned = islands_dev != nullptr;
    uint32_t* lrec = nullptr;
    // the state for the host: written by the emitting kernel into the page-locked stage
    SFState* hs = (SFState*)host_stage(ctx, "sf_state", sizeof(SFState));
    SFState* hs_dev = (SFState*)pinned_device_ptr(hs);
    auto island_sums = [&]() -> const IslandSum* {
        const size_t nl = size_t(ntiles) * CT_RT + rec.pool_cap;
        IslandSum* isum = wsp<IslandSum>(ctx, "sf_isum", nl);
        lrec = wsp<uint32_t>(ctx, "sf_lrec", nl);
        const SegIds sid{ tg, rmask, segoff };
        const unsigned rblocks = div_up(nl, 256);
        launch(ctx, "ccl_open_init", ccl_open_init, dim3(rblocks), dim3(256), 0, (const uint32_t*)P, sid, rec, isum, lrec,
            LabelCounts{ st, (const SelState*)ss, (unsigned long long)nl });
        launch(ctx, "ccl_records", ccl_records, dim3(rblocks), dim3(256), 0, (const uint32_t*)P, sid, rec, isum, lrec);
        return isum;
    };
    auto island_list = [&](const IslandSum* isum) {
        const size_t nl = size_t(ntiles) * CT_RT + rec.pool_cap;
        const unsigned rblocks = div_up(nl, 256);
        // the open islands' memory index (read by the emit only)
        launch(ctx, "ccl_open_idx", ccl_open_idx, dim3(rblocks), dim3(256), 0, (const uint32_t*)P,
            SegIds{ tg, rmask, segoff }, rec, (IslandSum*)isum);
        uint32_t* keep = wsp<uint32_t>(ctx, "sf_keep", nl);
        uint32_t* vcnt = wsp<uint32_t>(ctx, "sf_vcnt", 2 * size_t(rblocks));
        uint32_t* kcnt = vcnt + rblocks;
        disl = wsp<stp_island>(ctx, "sf_islands", nl);
        launch(ctx, "sf_island_flags", sf_island_flags, dim3(rblocks), dim3(256), 0, (const IslandSum*)isum,
            (const uint32_t*)lrec, (const IslandSum*)rec.sum, uint32_t(nl), (const SFState*)st, int(p.source_min_area), keep,
            vcnt, kcnt);
        launch(ctx, "sf_island_emit", sf_island_emit, dim3(rblocks), dim3(256), 0, (const IslandSum*)isum,
            (const uint32_t*)lrec, (const IslandSum*)rec.sum, (const uint32_t*)keep, (const uint32_t*)kcnt,
            (const uint32_t*)vcnt, st, rows, cols, p.analysis_n_sigma, pinned ? islands_dev : disl,
            pinned ? uint32_t(std::min<size_t>(cap, 0xFFFFFFFFu)) : uint32_t(nl), hs_dev);
    };
    // an ordinary caller buffer gets the island list through a stage: as many islands as
    // the last call found are copied with the state (a guess: any missing ones are
    // copied afterwards)
    const size_t spec = pinned ? 0 : std::min({ cap, ctx->sf_islands_hint, size_t(ntiles) * CT_RT + rec.pool_cap });
    stp_island* stage =
        pinned ? islands : (stp_island*)host_stage(ctx, "sf_islands", std::max<size_t>(1, spec) * sizeof(stp_island));
    // the island sums; then the label map on the copy stream while the island list is
    // built (and written or copied) and the state copied on the main stream; one host
    // wait covers both
    auto finish = [&]() {
        const IslandSum* isum = island_sums();
        hipStream_t cs = copy_stream(ctx);
        if (d_lab_out) {
            HIP_CHECK(hipEventRecord(ctx->copy_event, ctx->stream));
            HIP_CHECK(hipStreamWaitEvent(cs, ctx->copy_event, 0));
            launch_on(ctx, cs, "ccl_label_map", ccl_label_map, dim3(ntiles), dim3(LM_THREADS), 0, (const uint32_t*)P,
                SegIds{ tg, rmask, segoff }, (const SFState*)st, isum, (const uint32_t*)lrec, (const IslandSum*)rec.sum,
                int(p.generate_labelmap != 0), d_lab_out);
        }
        island_list(isum);
        if (!hs_dev)
            HIP_CHECK(hipMemcpyAsync(hs, st, sizeof(SFState), hipMemcpyDeviceToHost, ctx->stream));
        if (spec)
            HIP_CHECK(hipMemcpyAsync(stage, disl, spec * sizeof(stp_island), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipStreamSynchronize(cs));
    };
    rms_and_thresholds(pool_count);
    label(false);
    finish();
    if (hs->fallback) {
        // data the linear bins cannot resolve (heavy ties): radix selection on
        // the host-driven path, then the stages that depend on the median again
        radix_median<T>(ctx, x, n, st, sel);
        launch(ctx, "sf_fallback_clear", sf_fallback_clear, dim3(1), dim3(1), 0, ss);
        rms_and_thresholds(pool_count);
        label(false);
        finish();
    }
    if (hs->num_records > rec.pool_cap) {
        // more tile components than record slots: grow and label again (the
        // capacity is kept for later calls)
        records(size_t(hs->num_records) + size_t(hs->num_records) / 4);
        label(true);
        finish();
        if (hs->num_records > rec.pool_cap)
            throw Error(STP_ERR_INTERNAL, "component records overflow");
    }
    if (hs->skip)
        throw Error(STP_ERR_INTERNAL, "labelling incomplete");
    const uint32_t kp = hs->num_pos, kn = hs->num_neg;
    const size_t found = hs->num_islands;
    const size_t ncopy = std::min(found, cap);
    if (!pinned && ncopy > spec) {
        stage = (stp_island*)host_stage(ctx, "sf_islands", ncopy * sizeof(stp_island));
        HIP_CHECK(hipMemcpyAsync(stage, disl, ncopy * sizeof(stp_island), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    if (ncopy && !pinned)
        std::memcpy(islands, stage, ncopy * sizeof(stp_island));
    ctx->sf_islands_hint = std::max<size_t>(64, found + found / 8);
    *n_islands = found;
    if (stats) {
        stats->bg_level = hs->median;
        stats->rms_est = hs->rms;
        stats->mean = stats_valid ? hs->mean : 0.0;
        stats->sigma = stats_valid ? hs->sigma : 0.0;
        stats->num_labels_pos = kp;
        stats->num_labels_neg = kn;
        stats->num_valid_labels = hs->num_valid;
        stats->num_islands = uint32_t(found);
    }
    if (found > cap)
        throw Error(STP_ERR_CAPACITY, "island capacity too small: " + std::to_string(found) + " islands found");
}

} // namespace

void source_find_device(stp_ctx* ctx, const void* d_image, int precision, int rows, int cols, const stp_sourcefind_pars& p,
    int32_t* d_label_map, stp_island* islands, size_t cap, size_t* n_islands, stp_sourcefind_stats* stats)
{
    if (precision == STP_PRECISION_FP64)
        source_find_impl<double>(ctx, (const double*)d_image, rows, cols, p, d_label_map, islands, cap, n_islands, stats);
    else
        source_find_impl<float>(ctx, (const float*)d_image, rows, cols, p, d_label_map, islands, cap, n_islands, stats);
}

} // namespace stp_hip
