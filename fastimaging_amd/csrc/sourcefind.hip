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
//  labelling   union-find over the visual (fft-shifted) image with the
//              reference's ids: rank of each component's first pixel in the
//              scan (visual column, memory row) order (ccl.h:107-484,505-773);
//  islands     extrema (strict compare from the detection threshold, ties to the
//              smallest memory index), detection validity, label-map cleaning,
//              moments and counts (sourcefind.cpp:342-628); the island list and
//              the moments fit (:235-278, :691-717) are assembled on the host.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
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
    int pad;
    double num_sigma;
    // thresholds (sourcefind.cpp:298-301)
    double a_pos, a_neg, d_pos, d_neg;
    // results read back by the host
    unsigned int num_pos, num_neg;           // labels of each sign
    unsigned int num_valid, num_islands;     // detections; islands after the area cut
    int fallback;                            // exact median needs the radix selection
    int pad2;
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

// ------------------------------------------------------------------ sums
template <typename T>
__global__ __launch_bounds__(SWG) void sf_sum_kernel(const T* x, size_t n, double* part)
{
    __shared__ double lds[2 * CANON_LANES];
    const size_t base = size_t(blockIdx.x) * CANON_BLOCK;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t i = base + size_t(j) * CANON_LANES + threadIdx.x;
        if (i < n) {
            const double v = ldv(x, i);
            a = a + v;
            b = b + v * v;
        }
    }
    canon_tree(a, b, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

__global__ void sf_mean_sigma_kernel(const double* sums, unsigned long long n, SFState* st)
{
    const double mean = sums[0] / double(n);
    st->s1 = sums[0];
    st->s2 = sums[1];
    st->mean = mean;
    st->sigma = sqrt(sums[1] / double(n) - (mean * mean));
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
__global__ __launch_bounds__(SWG) void sf_binapprox_hist(const T* x, size_t n, const SFState* st, uint32_t* bh)
{
    __shared__ uint32_t h[BA_BINS + 1];
    if (!st->ba_active)
        return;
    for (int k = threadIdx.x; k < BA_BINS + 1; k += SWG)
        h[k] = 0;
    __syncthreads();
    const double scale = double(BA_BINS - 1) / (2 * st->sigma);
    const double left = st->mean - st->sigma;
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const int b = ba_bin(ldv(x, i), left, scale);
        if (b < 0)
            atomicAdd(&h[0], 1u);
        else if (b <= BA_BINS - 1)
            atomicAdd(&h[1 + b], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < BA_BINS + 1; k += SWG)
        bh[size_t(blockIdx.x) * (BA_BINS + 1) + k] = h[k];
}

// radix digit histogram (8 bits at `shift`) among elements whose key >> (shift + 8) == prefix
template <typename T>
__global__ __launch_bounds__(SWG) void sf_radix_hist(const T* x, size_t n, int shift, unsigned long long prefix, int use_prefix, uint32_t* bh)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const unsigned long long k = okey(ldv(x, i));
        if (use_prefix && (k >> (shift + 8)) != prefix)
            continue;
        atomicAdd(&h[(k >> shift) & 0xff], 1u);
    }
    __syncthreads();
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
    for (int b = b0; b < b1; b++)
        s += bh[size_t(b) * nb + k];
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
__global__ __launch_bounds__(SWG) void sf_gather(const T* x, size_t n, GatherSpec gs, unsigned long long* cand, unsigned* counter, unsigned cap)
{
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const double v = ldv(x, i);
        const bool take = (gs.shift >= 64) || ((okey(v) >> gs.shift) == gs.prefix);
        if (take) {
            const unsigned slot = atomicAdd(counter, 1u);
            if (slot < cap)
                cand[slot] = okey(v);
        }
    }
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

// Bins holding ranks ka <= kb of a histogram (1024 threads).  Writes the bin
// indices, the count before bin ia and the count through bin ib; -1 if absent.
__device__ void rank_bins(const unsigned long long* h, int nb, unsigned long long ka, unsigned long long kb, int* ia,
    int* ib, unsigned long long* before_a, unsigned long long* through_b)
{
    __shared__ unsigned long long part[1024];
    const int t = threadIdx.x;
    if (t == 0) {
        *ia = -1;
        *ib = -1;
    }
    const int per = (nb + 1023) / 1024;
    const int i0 = min(nb, t * per), i1 = min(nb, i0 + per);
    unsigned long long s = 0;
    for (int i = i0; i < i1; i++)
        s += h[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned long long v = (t >= off) ? part[t - off] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned long long cum = part[t] - s;
    for (int i = i0; i < i1; i++) {
        const unsigned long long c = cum + h[i];
        if (cum <= ka && c > ka) {
            *ia = i;
            *before_a = cum;
        }
        if (cum <= kb && c > kb) {
            *ib = i;
            *through_b = c;
        }
        cum = c;
    }
    __syncthreads();
}

__global__ void sel_init(SFState* st, SelState* ss, unsigned long long k_lo, unsigned long long k_hi, int odd, int force)
{
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

template <typename T>
__global__ __launch_bounds__(SWG) void sel_hist1(const T* x, size_t n, const SelState* ss, uint32_t* bh)
{
    __shared__ uint32_t h[LIN_BINS + 2];
    if (!ss->run1)
        return;
    const double lo = ss->lo, scale = ss->scale;
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += SWG)
        h[k] = 0;
    __syncthreads();
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG)
        atomicAdd(&h[lin_bin((ldv(x, i) - lo) * scale)], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += SWG)
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
__global__ __launch_bounds__(SWG) void sel_hist2(const T* x, size_t n, const SelState* ss, uint32_t* bh)
{
    __shared__ uint32_t h[LIN_BINS];
    if (!ss->run2)
        return;
    const double lo = ss->lo, scale = ss->scale, base2 = ss->base2, scale2 = ss->scale2;
    const int b1lo = ss->b1_lo, b1hi = ss->b1_hi;
    for (int k = threadIdx.x; k < LIN_BINS; k += SWG)
        h[k] = 0;
    __syncthreads();
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const double t = (ldv(x, i) - lo) * scale;
        const int b = lin_bin(t);
        if (b >= b1lo && b <= b1hi)
            atomicAdd(&h[sub_bin(t, base2, scale2)], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < LIN_BINS; k += SWG)
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
__global__ __launch_bounds__(SWG) void sel_gather(const T* x, size_t n, const SelState* ss, unsigned long long* cand,
    unsigned* counter)
{
    if (!ss->gather)
        return;
    const double lo = ss->lo, scale = ss->scale, base2 = ss->base2, scale2 = ss->scale2;
    const int b1lo = ss->b1_lo, b1hi = ss->b1_hi, b2lo = ss->b2_lo, b2hi = ss->b2_hi, lvl = ss->level;
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const double v = ldv(x, i);
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
    }
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
__global__ void sf_clip_init(SFState* st, unsigned long long n, int iters, double num_sigma)
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
}

template <typename T>
__global__ __launch_bounds__(SWG) void sf_clip_kernel(const T* x, size_t n, const SFState* st, double* part, uint32_t* removed)
{
    __shared__ double lds[2 * CANON_LANES];
    __shared__ uint32_t wc[SWG / 64];
    if (st->done)
        return;
    const double median = st->median, upper = st->upper, prev = st->prev_upper;
    const size_t base = size_t(blockIdx.x) * CANON_BLOCK;
    double a = 0.0, b = 0.0;
    unsigned cnt = 0;
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t i = base + size_t(j) * CANON_LANES + threadIdx.x;
        if (i < n) {
            const double v = ldv(x, i) - median;
            const double av = fabs(v);
            const bool rem = (av > upper) && !(av > prev);
            a = a + (rem ? v : 0.0);
            b = b + (rem ? v * v : 0.0);
            cnt += rem ? 1u : 0u;
        }
    }
    canon_tree(a, b, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
    for (int off = 32; off >= 1; off >>= 1)
        cnt += __shfl_down(cnt, off, 64);
    if ((threadIdx.x & 63) == 0)
        wc[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int w = 0; w < SWG / 64; w++)
            c += wc[w];
        removed[blockIdx.x] = c;
    }
}

// One clipping step from the block partials (removed: one count per block).
__global__ __launch_bounds__(256) void sf_clip_update(SFState* st, const double* sums, const uint32_t* removed, int nblk)
{
    __shared__ unsigned long long red[256];
    if (st->done)
        return;
    unsigned long long c = 0;
    for (int b = threadIdx.x; b < nblk; b += 256)
        c += removed[b];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0)
        return;
    const unsigned long long rem = red[0];
    if (rem == 0) {
        st->done = 1;
        return;
    }
    st->valid_n -= rem;
    st->d1 -= sums[0];
    st->d2 -= sums[1];
    const double vn = double(st->valid_n);
    st->clip_sigma = sqrt(st->d2 / vn - (st->d1 / vn) * (st->d1 / vn));
    st->prev_upper = st->upper;
    st->upper = st->num_sigma * st->clip_sigma;
    st->iter++;
    if (st->iter >= st->iters)
        st->done = 1;
}

__global__ void sf_thresholds(SFState* st, int estimate, double rms_in, double det, double ana)
{
    const double rms = estimate ? st->clip_sigma : rms_in;
    st->rms = rms;
    const double bg = st->median;
    st->a_pos = bg + ana * rms;
    st->a_neg = bg - ana * rms;
    st->d_pos = bg + det * rms;
    st->d_neg = bg - det * rms;
}

// ------------------------------------------------------------------ labelling
// key of memory pixel (r, mc): visual column cv = (mc + cols/2) % cols; key = cv * rows + r
template <typename T>
__global__ __launch_bounds__(SWG) void ccl_classify(const T* x, int rows, int cols, const SFState* st, int find_neg,
    uint8_t* cls, uint32_t* P)
{
    const size_t n = size_t(rows) * cols;
    const size_t m = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (m >= n)
        return;
    const int r = int(m % rows), mc = int(m / rows);
    const int cv = (mc + cols / 2) % cols;
    const size_t k = size_t(cv) * rows + r;
    const double v = ldv(x, m);
    uint8_t c = 0;
    if (v > st->a_pos)
        c = 1;
    else if (find_neg && v < st->a_neg)
        c = 2;
    cls[k] = c;
    if (c)
        P[k] = uint32_t(k);
}

__device__ __forceinline__ uint32_t uf_find(const uint32_t* P, uint32_t x)
{
    uint32_t p = P[x];
    while (p != x) {
        x = p;
        p = P[x];
    }
    return x;
}

// Links the trees of a and b; the smaller key becomes the root (lock-free, atomicMin).
__device__ void uf_union(uint32_t* P, uint32_t a, uint32_t b)
{
    bool done;
    do {
        a = uf_find(P, a);
        b = uf_find(P, b);
        if (a < b) {
            const uint32_t old = atomicMin(&P[b], a);
            done = (old == b);
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(&P[a], b);
            done = (old == a);
            a = old;
        } else {
            done = true;
        }
    } while (!done);
}

__global__ __launch_bounds__(SWG) void ccl_merge(const uint8_t* cls, uint32_t* P, int rows, int cols, int conn4)
{
    const size_t n = size_t(rows) * cols;
    const size_t k = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (k >= n)
        return;
    const uint8_t c = cls[k];
    if (!c)
        return;
    const int r = int(k % rows), cv = int(k / rows);
    const int rshift = rows / 2;
    const int vr = (r + rshift) % rows;
    auto key_of = [&](int ncv, int nvr) -> uint32_t { return uint32_t(size_t(ncv) * rows + (nvr + rows - rshift) % rows); };
    // backward neighbours in the visual image, no wrap: left, up, up-left, down-left
    if (cv > 0) {
        const uint32_t nk = key_of(cv - 1, vr);
        if (cls[nk] == c)
            uf_union(P, uint32_t(k), nk);
    }
    if (vr > 0) {
        const uint32_t nk = key_of(cv, vr - 1);
        if (cls[nk] == c)
            uf_union(P, uint32_t(k), nk);
    }
    if (!conn4 && cv > 0) {
        if (vr > 0) {
            const uint32_t nk = key_of(cv - 1, vr - 1);
            if (cls[nk] == c)
                uf_union(P, uint32_t(k), nk);
        }
        if (vr < rows - 1) {
            const uint32_t nk = key_of(cv - 1, vr + 1);
            if (cls[nk] == c)
                uf_union(P, uint32_t(k), nk);
        }
    }
}

__global__ __launch_bounds__(SWG) void ccl_flatten(const uint8_t* cls, uint32_t* P, size_t n)
{
    const size_t k = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (k < n && cls[k])
        P[k] = uf_find(P, uint32_t(k));
}

// roots per block of CANON_BLOCK keys (thread t owns keys [t*16, t*16+16) of the block)
__global__ __launch_bounds__(SWG) void ccl_root_count(const uint8_t* cls, const uint32_t* P, size_t n, uint32_t* cpos, uint32_t* cneg)
{
    __shared__ uint32_t sp[SWG], sn[SWG];
    const size_t base = size_t(blockIdx.x) * CANON_BLOCK + size_t(threadIdx.x) * CANON_PER_LANE;
    uint32_t p = 0, q = 0;
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t k = base + j;
        if (k < n && cls[k] && P[k] == k) {
            if (cls[k] == 1)
                p++;
            else
                q++;
        }
    }
    sp[threadIdx.x] = p;
    sn[threadIdx.x] = q;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int t = 0; t < SWG; t++) {
            a += sp[t];
            b += sn[t];
        }
        cpos[blockIdx.x] = a;
        cneg[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(SWG) void ccl_root_ids(const uint8_t* cls, const uint32_t* P, size_t n, const uint32_t* opos,
    const uint32_t* oneg, int32_t* ID)
{
    __shared__ uint32_t sp[SWG], sn[SWG];
    const size_t base = size_t(blockIdx.x) * CANON_BLOCK + size_t(threadIdx.x) * CANON_PER_LANE;
    uint32_t p = 0, q = 0;
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t k = base + j;
        if (k < n && cls[k] && P[k] == k) {
            if (cls[k] == 1)
                p++;
            else
                q++;
        }
    }
    sp[threadIdx.x] = p;
    sn[threadIdx.x] = q;
    __syncthreads();
    // exclusive prefix over the threads of the block (serial in LDS by thread 0 would be slow; Hillis-Steele)
    for (int off = 1; off < SWG; off <<= 1) {
        const uint32_t a = threadIdx.x >= off ? sp[threadIdx.x - off] : 0u;
        const uint32_t b = threadIdx.x >= off ? sn[threadIdx.x - off] : 0u;
        __syncthreads();
        sp[threadIdx.x] += a;
        sn[threadIdx.x] += b;
        __syncthreads();
    }
    uint32_t rp = opos[blockIdx.x] + sp[threadIdx.x] - p;
    uint32_t rn = oneg[blockIdx.x] + sn[threadIdx.x] - q;
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const size_t k = base + j;
        if (k < n && cls[k] && P[k] == k) {
            if (cls[k] == 1)
                ID[k] = int32_t(++rp);
            else
                ID[k] = -int32_t(++rn);
        }
    }
}

__global__ __launch_bounds__(SWG) void ccl_label_out(const uint8_t* cls, const uint32_t* P, const int32_t* ID, int rows, int cols,
    int32_t* lab)
{
    const size_t n = size_t(rows) * cols;
    const size_t m = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (m >= n)
        return;
    const int r = int(m % rows), mc = int(m / rows);
    const size_t k = size_t((mc + cols / 2) % cols) * rows + r;
    lab[m] = cls[k] ? ID[P[k]] : 0;
}

// ------------------------------------------------------------------ islands
template <typename T>
__global__ __launch_bounds__(SWG) void sf_extrema(const T* x, const int32_t* lab, size_t n, unsigned long long* ep, unsigned long long* en)
{
    const size_t m = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (m >= n)
        return;
    const int32_t l = lab[m];
    if (l > 0)
        atomicMax(&ep[l - 1], okey(ldv(x, m)));
    else if (l < 0)
        atomicMin(&en[-l - 1], okey(ldv(x, m)));
}

// Memory index of each island's extremum: the smallest index holding the extremal key.
template <typename T>
__global__ __launch_bounds__(SWG) void sf_extrema_idx(const T* x, const int32_t* lab, size_t n, const unsigned long long* ep,
    const unsigned long long* en, unsigned long long* ip, unsigned long long* in_)
{
    const size_t m = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (m >= n)
        return;
    const int32_t l = lab[m];
    if (l > 0) {
        if (okey(ldv(x, m)) == ep[l - 1])
            atomicMin(&ip[l - 1], (unsigned long long)m);
    } else if (l < 0) {
        if (okey(ldv(x, m)) == en[-l - 1])
            atomicMin(&in_[-l - 1], (unsigned long long)m);
    }
}

__global__ void sf_init_labels(unsigned long long* ep, unsigned long long* en, unsigned long long* ip, unsigned long long* in_,
    int* cp, int* cn, double* mp, double* mn, uint32_t kp, uint32_t kn, const SFState* st)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < kp) {
        ep[i] = okey(st->d_pos);
        ip[i] = ~0ull;
        cp[i] = 0;
        for (int q = 0; q < 6; q++)
            mp[6 * i + q] = 0.0;
    }
    if (i < kn) {
        en[i] = okey(st->d_neg);
        in_[i] = ~0ull;
        cn[i] = 0;
        for (int q = 0; q < 6; q++)
            mn[6 * i + q] = 0.0;
    }
}

template <typename T>
__global__ __launch_bounds__(SWG) void sf_moments(const T* x, int32_t* lab, int rows, int cols, const SFState* st,
    const unsigned long long* ep, const unsigned long long* en, int gen_map, int* cp, int* cn, double* mp, double* mn)
{
    const size_t n = size_t(rows) * cols;
    const size_t m = size_t(blockIdx.x) * SWG + threadIdx.x;
    if (m >= n)
        return;
    const int32_t l = lab[m];
    if (!l)
        return;
    // detection validity: extremum strictly beyond the detection threshold (sourcefind.cpp:417-436)
    const bool valid = (l > 0) ? (okey_inv(ep[l - 1]) > st->d_pos) : (okey_inv(en[-l - 1]) < st->d_neg);
    if (!valid) {
        if (gen_map)
            lab[m] = 0;
        return;
    }
    const int i = int(m / rows), j = int(m % rows);
    const int h_shift = cols / 2, v_shift = rows / 2;
    const double val = ldv(x, m);
    const double y_idx = j < v_shift ? double(j + v_shift) : double(j - v_shift);
    const double x_idx = i < h_shift ? double(i + h_shift) : double(i - h_shift);
    const double xb = x_idx * val, yb = y_idx * val;
    const double xx = xb * x_idx, yy = yb * y_idx, xy = xb * y_idx;
    double* mm = (l > 0) ? &mp[6 * size_t(l - 1)] : &mn[6 * size_t(-l - 1)];
    atomicAdd((l > 0) ? &cp[l - 1] : &cn[-l - 1], 1);
    unsafeAtomicAdd(&mm[0], xb);
    unsafeAtomicAdd(&mm[1], yb);
    unsafeAtomicAdd(&mm[2], xx);
    unsafeAtomicAdd(&mm[3], yy);
    unsafeAtomicAdd(&mm[4], xy);
    unsafeAtomicAdd(&mm[5], val);
}

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
__global__ __launch_bounds__(256) void sf_island_flags(const unsigned long long* ep, const unsigned long long* en,
    const int* cp, const int* cn, uint32_t kp, uint32_t kn, const SFState* st, int min_area, uint32_t* keep, uint32_t* valid)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= kp + kn)
        return;
    const bool pos = g < kp;
    const uint32_t i = pos ? g : g - kp;
    const double ext = okey_inv(pos ? ep[i] : en[i]);
    const bool v = pos ? (ext > st->d_pos) : (ext < st->d_neg);
    const int c = pos ? cp[i] : cn[i];
    valid[g] = v ? 1u : 0u;
    keep[g] = (v && c >= min_area) ? 1u : 0u;
}

// One island record per kept label at its rank among the kept labels.
__global__ __launch_bounds__(256) void sf_island_emit(const unsigned long long* ep, const unsigned long long* en,
    const unsigned long long* ip, const unsigned long long* in_, const int* cp, const int* cn, const double* mp,
    const double* mn, uint32_t kp, uint32_t kn, const uint32_t* keep, const uint32_t* offs, const SFState* st, int rows,
    int cols, double analysis_n_sigma, stp_island* out)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= kp + kn || !keep[g])
        return;
    const bool pos = g < kp;
    const uint32_t i = pos ? g : g - kp;
    const size_t li = size_t(pos ? ip[i] : in_[i]);
    const int r = int(li % size_t(rows)), c = int(li / size_t(rows));
    const int h_shift = cols / 2, v_shift = rows / 2;
    stp_island isl{};
    isl.label_idx = pos ? int32_t(i + 1) : -int32_t(i + 1);
    isl.sign = pos ? 1 : -1;
    isl.extremum_y_idx = r < v_shift ? r + v_shift : r - v_shift;
    isl.extremum_x_idx = c < h_shift ? c + h_shift : c - h_shift;
    isl.num_samples = pos ? cp[i] : cn[i];
    isl.extremum_val = okey_inv(pos ? ep[i] : en[i]);
    // moments normalised by the flux sum (sourcefind.cpp:594-613)
    const double* m = pos ? &mp[6 * size_t(i)] : &mn[6 * size_t(i)];
    const double xb = m[0] / m[5], yb = m[1] / m[5];
    moments_fit(isl, xb, yb, m[2] / m[5] - xb * xb, m[3] / m[5] - yb * yb, m[4] / m[5] - xb * yb, st->rms,
        analysis_n_sigma);
    out[offs[g]] = isl;
}

__global__ void sf_label_counts(SFState* st, const uint32_t* pos_total, const uint32_t* neg_total, const SelState* ss)
{
    st->num_pos = *pos_total;
    st->num_neg = *neg_total;
    st->fallback = ss->fallback;
}

__global__ void sf_island_counts(SFState* st, const uint32_t* keep_total, const uint32_t* valid_total)
{
    st->num_islands = keep_total ? *keep_total : 0u;
    st->num_valid = valid_total ? *valid_total : 0u;
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

unsigned hist_blocks(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>(512, (n + 16383) / 16384))); }


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
        launch(ctx, "sf_radix_hist", sf_radix_hist<T>, dim3(nb), dim3(SWG), 0, x, n, shift, prefix, int(shift < 56), rbh);
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
    launch(ctx, "sf_gather", sf_gather<T>, dim3(nb), dim3(SWG), 0, x, n, gs, cand, counter, unsigned(SORT_CAP));
    launch(ctx, "sf_sort_select", sf_sort_select, dim3(1), dim3(1024), 0, (const unsigned long long*)cand,
        (const unsigned*)counter, int(k - before), int(k - before), tmp);
    HIP_CHECK(hipMemcpyAsync(d_out, tmp, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

// Enqueues the device-driven exact median search (SelState).  It runs when
// st->need_exact is set on the device (force sets it), reads st->mean and
// st->sigma, and sets st->median; ss->fallback reports data the bins cannot
// resolve.  No host round trip.
template <typename T>
void enqueue_exact_median(stp_ctx* ctx, const T* x, size_t n, SFState* st, SelState* ss, int force)
{
    const unsigned nb = hist_blocks(n);
    uint32_t* bh = wsp<uint32_t>(ctx, "sf_bh", size_t(nb) * (LIN_BINS + 2));
    unsigned long long* hist = wsp<unsigned long long>(ctx, "sf_hist", LIN_BINS + 2);
    unsigned long long* cand = wsp<unsigned long long>(ctx, "sf_cand", SORT_CAP);
    unsigned* counter = wsp<unsigned>(ctx, "sf_counter", 1);
    const unsigned long long k_hi = n / 2, k_lo = (n & 1) ? n / 2 : n / 2 - 1;
    const unsigned slices = std::min(16u, nb);
    launch(ctx, "sel_init", sel_init, dim3(1), dim3(1), 0, st, ss, k_lo, k_hi, int(n & 1), force);
    launch(ctx, "sel_hist1", sel_hist1<T>, dim3(nb), dim3(SWG), 0, x, n, (const SelState*)ss, bh);
    fill_u32(ctx, (uint32_t*)hist, 0u, 2 * size_t(LIN_BINS + 2));
    launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(LIN_BINS + 2, 256), slices), dim3(256), 0,
        (const uint32_t*)bh, int(nb), LIN_BINS + 2, hist, (const int*)&ss->run1);
    launch(ctx, "sel_pick1", sel_pick1, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)hist);
    launch(ctx, "sel_hist2", sel_hist2<T>, dim3(nb), dim3(SWG), 0, x, n, (const SelState*)ss, bh);
    fill_u32(ctx, (uint32_t*)hist, 0u, 2 * size_t(LIN_BINS));
    launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(LIN_BINS, 256), slices), dim3(256), 0,
        (const uint32_t*)bh, int(nb), LIN_BINS, hist, (const int*)&ss->run2);
    launch(ctx, "sel_pick2", sel_pick2, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)hist);
    fill_u32(ctx, counter, 0u, 1);
    launch(ctx, "sel_gather", sel_gather<T>, dim3(nb), dim3(SWG), 0, x, n, (const SelState*)ss, cand, counter);
    launch(ctx, "sel_sort", sel_sort, dim3(1), dim3(1024), 0, ss, (const unsigned long long*)cand, (const unsigned*)counter, st);
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
    STP_REQUIRE(n < (size_t(1) << 31), "image too large for 32-bit labels");
    if (p.gaussian_fitting)
        throw Error(STP_ERR_UNSUPPORTED, "Gaussian least-squares fitting is not implemented");
    STP_REQUIRE(p.median_method >= STP_MEDIAN_ZEROMEDIAN && p.median_method <= STP_MEDIAN_NTHELEMENT, "unknown median method");

    SFState* st = wsp<SFState>(ctx, "sf_state", 1);
    double* sel = wsp<double>(ctx, "sf_sel", 2);
    const size_t nblk = (n + CANON_BLOCK - 1) / CANON_BLOCK;
    double* part = wsp<double>(ctx, "sf_part", 2 * nblk + 2);
    double* scr = wsp<double>(ctx, "sf_scr", canon_scratch_doubles(nblk));
    double* sums = wsp<double>(ctx, "sf_sums", 2);
    const bool need_rms = !(std::fabs(p.rms_est) > 0.0);
    const int mm = p.median_method;
    bool stats_valid = false;
    auto compute_mean_sigma = [&]() {
        launch(ctx, "sf_sum", sf_sum_kernel<T>, dim3(unsigned(nblk)), dim3(SWG), 0, x, n, part);
        canon_reduce_pairs(ctx, part, nblk, scr, sums);
        launch(ctx, "sf_stats", sf_mean_sigma_kernel, dim3(1), dim3(1), 0, (const double*)sums, (unsigned long long)n, st);
        stats_valid = true;
    };
    SelState* ss = wsp<SelState>(ctx, "sf_selstate", 1);
    fill_f64(ctx, (double*)st, 0.0, sizeof(SFState) / sizeof(double));
    fill_f64(ctx, (double*)ss, 0.0, sizeof(SelState) / sizeof(double));
    // median (sourcefind.cpp:182-197), enqueued without host round trips: the
    // BINAPPROX pick and the exact search are gated on the device state
    if (mm == STP_MEDIAN_BINAPPROX) {
        compute_mean_sigma();
        launch(ctx, "sf_ba_gate", sf_ba_gate, dim3(1), dim3(1), 0, st);
        const unsigned nb = hist_blocks(n);
        uint32_t* bh = wsp<uint32_t>(ctx, "sf_babh", size_t(nb) * (BA_BINS + 1));
        unsigned long long* hist = wsp<unsigned long long>(ctx, "sf_bahist", BA_BINS + 1);
        launch(ctx, "sf_binapprox_hist", sf_binapprox_hist<T>, dim3(nb), dim3(SWG), 0, x, n, (const SFState*)st, bh);
        fill_u32(ctx, (uint32_t*)hist, 0u, 2 * size_t(BA_BINS + 1));
        launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(BA_BINS + 1, 256), std::min(64u, nb)), dim3(256), 0,
            (const uint32_t*)bh, int(nb), BA_BINS + 1, hist, (const int*)&st->ba_active);
        launch(ctx, "sf_binapprox_pick", sf_binapprox_pick, dim3(1), dim3(1), 0, (const unsigned long long*)hist,
            (unsigned long long)n, st);
        enqueue_exact_median<T>(ctx, x, n, st, ss, 0);
    } else if (mm == STP_MEDIAN_BINMEDIAN || mm == STP_MEDIAN_NTHELEMENT) {
        // mean and sigma seed the selection's linear bins (and estimate_rms,
        // sourcefind.cpp:30-36)
        compute_mean_sigma();
        enqueue_exact_median<T>(ctx, x, n, st, ss, 1);
    }
    // ZEROMEDIAN: the state's median stays 0
    if (need_rms && !stats_valid)
        compute_mean_sigma();

    // RMS (sourcefind.cpp:206) and thresholds (:298-301)
    auto rms_and_thresholds = [&]() {
        if (need_rms) {
            const int iters = std::max(0, p.sigma_clip_iters);
            uint32_t* removed = wsp<uint32_t>(ctx, "sf_removed", nblk);
            double* cpart = wsp<double>(ctx, "sf_cpart", 2 * nblk + 2);
            double* csums = wsp<double>(ctx, "sf_csums", 2);
            launch(ctx, "sf_clip_init", sf_clip_init, dim3(1), dim3(1), 0, st, (unsigned long long)n, iters, 3.0);
            for (int it = 0; it < iters; it++) {
                launch(ctx, "sf_clip", sf_clip_kernel<T>, dim3(unsigned(nblk)), dim3(SWG), 0, x, n, (const SFState*)st, cpart,
                    removed);
                canon_reduce_pairs(ctx, cpart, nblk, scr, csums);
                launch(ctx, "sf_clip_update", sf_clip_update, dim3(1), dim3(256), 0, st, (const double*)csums,
                    (const uint32_t*)removed, int(nblk));
            }
        }
        launch(ctx, "sf_thresholds", sf_thresholds, dim3(1), dim3(1), 0, st, int(need_rms), p.rms_est,
            p.detection_n_sigma, p.analysis_n_sigma);
    };

    // connected components (ccl.h), ids in the reference's scan order
    uint8_t* cls = wsp<uint8_t>(ctx, "ccl_cls", n);
    uint32_t* P = wsp<uint32_t>(ctx, "ccl_parent", n);
    int32_t* ID = wsp<int32_t>(ctx, "ccl_id", n);
    int32_t* lab = d_lab_out ? d_lab_out : wsp<int32_t>(ctx, "ccl_label", n);
    uint32_t* cpos = wsp<uint32_t>(ctx, "ccl_cpos", nblk);
    uint32_t* cneg = wsp<uint32_t>(ctx, "ccl_cneg", nblk);
    uint32_t* opos = wsp<uint32_t>(ctx, "ccl_opos", nblk + 1);
    uint32_t* oneg = wsp<uint32_t>(ctx, "ccl_oneg", nblk + 1);
    uint32_t* cur = wsp<uint32_t>(ctx, "ccl_cur", nblk);
    auto label = [&]() {
    launch(ctx, "ccl_classify", ccl_classify<T>, dim3(div_up(n, SWG)), dim3(SWG), 0, x, rows, cols, (const SFState*)st,
        int(p.find_negative_sources != 0), cls, P);
    launch(ctx, "ccl_merge", ccl_merge, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, P, rows, cols,
        int(p.ccl_4connectivity != 0));
    launch(ctx, "ccl_flatten", ccl_flatten, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, P, n);
    launch(ctx, "ccl_root_count", ccl_root_count, dim3(unsigned(nblk)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P,
        n, cpos, cneg);
    exclusive_scan_u32(ctx, cpos, int(nblk), opos, cur);
    exclusive_scan_u32(ctx, cneg, int(nblk), oneg, cur);
    launch(ctx, "ccl_root_ids", ccl_root_ids, dim3(unsigned(nblk)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P, n,
        (const uint32_t*)opos, (const uint32_t*)oneg, ID);
    launch(ctx, "ccl_label_out", ccl_label_out, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P,
        (const int32_t*)ID, rows, cols, lab);
    };
    SFState* hs = (SFState*)host_stage(ctx, "sf_state", sizeof(SFState));
    auto read_state = [&]() {
        HIP_CHECK(hipMemcpyAsync(hs, st, sizeof(SFState), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
    };
    rms_and_thresholds();
    label();
    launch(ctx, "sf_counts", sf_label_counts, dim3(1), dim3(1), 0, st, (const uint32_t*)(opos + nblk),
        (const uint32_t*)(oneg + nblk), (const SelState*)ss);
    read_state();
    if (hs->fallback) {
        // data the linear bins cannot resolve (heavy ties): radix selection on
        // the host-driven path, then the stages that depend on the median again
        radix_median<T>(ctx, x, n, st, sel);
        launch(ctx, "sf_fallback_clear", sf_fallback_clear, dim3(1), dim3(1), 0, ss);
        rms_and_thresholds();
        label();
        launch(ctx, "sf_counts", sf_label_counts, dim3(1), dim3(1), 0, st, (const uint32_t*)(opos + nblk),
            (const uint32_t*)(oneg + nblk), (const SelState*)ss);
        read_state();
    }
    const uint32_t kp = hs->num_pos, kn = hs->num_neg;

    // extrema, validity, moments (sourcefind.cpp:337-628)
    unsigned long long* ep = wsp<unsigned long long>(ctx, "sf_ep", kp);
    unsigned long long* en = wsp<unsigned long long>(ctx, "sf_en", kn);
    unsigned long long* ip = wsp<unsigned long long>(ctx, "sf_ip", kp);
    unsigned long long* in_ = wsp<unsigned long long>(ctx, "sf_in", kn);
    int* cp = wsp<int>(ctx, "sf_cp", kp);
    int* cn = wsp<int>(ctx, "sf_cn", kn);
    double* mp = wsp<double>(ctx, "sf_mp", 6 * size_t(kp));
    double* mn = wsp<double>(ctx, "sf_mn", 6 * size_t(kn));
    launch(ctx, "sf_init_labels", sf_init_labels, dim3(div_up(std::max(kp, kn), 256)), dim3(256), 0, ep, en, ip, in_, cp, cn,
        mp, mn, kp, kn, (const SFState*)st);
    launch(ctx, "sf_extrema", sf_extrema<T>, dim3(div_up(n, SWG)), dim3(SWG), 0, x, (const int32_t*)lab, n, ep, en);
    launch(ctx, "sf_extrema_idx", sf_extrema_idx<T>, dim3(div_up(n, SWG)), dim3(SWG), 0, x, (const int32_t*)lab, n,
        (const unsigned long long*)ep, (const unsigned long long*)en, ip, in_);
    launch(ctx, "sf_moments", sf_moments<T>, dim3(div_up(n, SWG)), dim3(SWG), 0, x, lab, rows, cols, (const SFState*)st,
        (const unsigned long long*)ep, (const unsigned long long*)en, int(p.generate_labelmap != 0), cp, cn, mp, mn);

    // island list: validity, area cut, compaction and moments fit on the device
    const size_t nl = size_t(kp) + size_t(kn);
    uint32_t* keep = wsp<uint32_t>(ctx, "sf_keep", nl);
    uint32_t* valid = wsp<uint32_t>(ctx, "sf_valid", nl);
    uint32_t* koff = wsp<uint32_t>(ctx, "sf_koff", nl + 1);
    uint32_t* voff = wsp<uint32_t>(ctx, "sf_voff", nl + 1);
    stp_island* disl = wsp<stp_island>(ctx, "sf_islands", nl);
    if (nl) {
        launch(ctx, "sf_island_flags", sf_island_flags, dim3(unsigned(div_up(nl, 256))), dim3(256), 0,
            (const unsigned long long*)ep, (const unsigned long long*)en, (const int*)cp, (const int*)cn, kp, kn,
            (const SFState*)st, int(p.source_min_area), keep, valid);
        scan_u32(ctx, keep, nl, koff);
        scan_u32(ctx, valid, nl, voff);
        launch(ctx, "sf_island_emit", sf_island_emit, dim3(unsigned(div_up(nl, 256))), dim3(256), 0,
            (const unsigned long long*)ep, (const unsigned long long*)en, (const unsigned long long*)ip,
            (const unsigned long long*)in_, (const int*)cp, (const int*)cn, (const double*)mp, (const double*)mn, kp, kn,
            (const uint32_t*)keep, (const uint32_t*)koff, (const SFState*)st, rows, cols, p.analysis_n_sigma, disl);
        launch(ctx, "sf_island_counts", sf_island_counts, dim3(1), dim3(1), 0, st, (const uint32_t*)(koff + nl),
            (const uint32_t*)(voff + nl));
    } else {
        launch(ctx, "sf_island_counts", sf_island_counts, dim3(1), dim3(1), 0, st, (const uint32_t*)nullptr,
            (const uint32_t*)nullptr);
    }
    read_state();
    const size_t found = hs->num_islands;
    const size_t ncopy = std::min(found, cap);
    if (ncopy) {
        stp_island* stage = (stp_island*)host_stage(ctx, "sf_islands", ncopy * sizeof(stp_island));
        HIP_CHECK(hipMemcpyAsync(stage, disl, ncopy * sizeof(stp_island), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        std::memcpy(islands, stage, ncopy * sizeof(stp_island));
    }
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
