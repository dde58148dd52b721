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
    int iter, iters, done, pad;
    double num_sigma;
    // thresholds (sourcefind.cpp:298-301)
    double a_pos, a_neg, d_pos, d_neg;
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

// linear bins over [lo, lo + nb/scale): index 0 = below, 1..nb = bins, nb+1 = above (or NaN)
template <typename T>
__global__ __launch_bounds__(SWG) void sf_lin_hist(const T* x, size_t n, double lo, double scale, uint32_t* bh)
{
    __shared__ uint32_t h[LIN_BINS + 2];
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += SWG)
        h[k] = 0;
    __syncthreads();
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const double t = (ldv(x, i) - lo) * scale;
        int b;
        if (t < 0.0)
            b = 0;
        else if (t < double(LIN_BINS))
            b = 1 + int(t);
        else
            b = LIN_BINS + 1;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < LIN_BINS + 2; k += SWG)
        bh[size_t(blockIdx.x) * (LIN_BINS + 2) + k] = h[k];
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

__global__ void sf_hist_reduce(const uint32_t* bh, int nblk, int nb, unsigned long long* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb)
        return;
    unsigned long long s = 0;
    for (int b = 0; b < nblk; b++)
        s += bh[size_t(b) * nb + k];
    out[k] = s;
}

// BINAPPROX median from the histogram (matrix_math.cpp:406-452); found = 0 -> exact fallback.
__global__ void sf_binapprox_pick(const unsigned long long* h, unsigned long long n, SFState* st, int* found)
{
    const int N = BA_BINS - 1;
    const double scale = double(N) / (2 * st->sigma);
    const double left = st->mean - st->sigma;
    unsigned long long count = h[0];
    *found = 0;
    if (n & 1) {
        const unsigned long long k = (n + 1) / 2;
        for (unsigned i = 0; i < unsigned(N + 1); i++) {
            count += h[1 + i];
            if (count >= k) {
                st->median = double(i + 0.5) / scale + left;
                *found = 1;
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
                *found = 1;
                return;
            }
        }
    }
}

// ------------------------------------------------------------------ candidate gather + sort
struct GatherSpec {
    int mode; // 0: linear bins [b_lo, b_hi] (1-based as in sf_lin_hist), 1: key prefix
    double lo, scale;
    int b_lo, b_hi;
    int shift; // prefix: key >> shift == prefix
    unsigned long long prefix;
};

template <typename T>
__global__ __launch_bounds__(SWG) void sf_gather(const T* x, size_t n, GatherSpec gs, unsigned long long* cand, unsigned* counter, unsigned cap)
{
    for (size_t i = size_t(blockIdx.x) * SWG + threadIdx.x; i < n; i += size_t(gridDim.x) * SWG) {
        const double v = ldv(x, i);
        bool take;
        if (gs.mode == 0) {
            const double t = (v - gs.lo) * gs.scale;
            int b;
            if (t < 0.0)
                b = 0;
            else if (t < double(LIN_BINS))
                b = 1 + int(t);
            else
                b = LIN_BINS + 1;
            take = b >= gs.b_lo && b <= gs.b_hi;
        } else {
            // shift 64: no prefix selected yet, every element is a candidate
            take = (gs.shift >= 64) || ((okey(v) >> gs.shift) == gs.prefix);
        }
        if (take) {
            const unsigned slot = atomicAdd(counter, 1u);
            if (slot < cap)
                cand[slot] = okey(v);
        }
    }
}

// Sorts count (<= SORT_CAP) keys in LDS and writes the values of ranks r0 and r1
// (0-based within the candidates) to out[0], out[1].
__global__ __launch_bounds__(1024) void sf_sort_select(const unsigned long long* cand, const unsigned* count, int r0, int r1, double* out)
{
    __shared__ unsigned long long s[SORT_CAP];
    const unsigned c = min(*count, unsigned(SORT_CAP));
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
__global__ __launch_bounds__(SWG) void sf_clip_kernel(const T* x, size_t n, const SFState* st, double* part, unsigned long long* removed)
{
    __shared__ double lds[2 * CANON_LANES];
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
    if ((threadIdx.x & 63) == 0 && cnt)
        atomicAdd(&removed[st->iter], (unsigned long long)cnt);
}

__global__ void sf_clip_update(SFState* st, const double* sums, const unsigned long long* removed)
{
    if (st->done)
        return;
    const unsigned long long rem = removed[st->iter];
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
        launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(1), dim3(256), 0, (const uint32_t*)rbh, int(nb), 256, rhist);
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
    gs.mode = 1;
    gs.shift = shift;
    gs.prefix = prefix;
    fill_u32(ctx, counter, 0u, 1);
    launch(ctx, "sf_gather", sf_gather<T>, dim3(nb), dim3(SWG), 0, x, n, gs, cand, counter, unsigned(SORT_CAP));
    launch(ctx, "sf_sort_select", sf_sort_select, dim3(1), dim3(1024), 0, (const unsigned long long*)cand,
        (const unsigned*)counter, int(k - before), int(k - before), tmp);
    HIP_CHECK(hipMemcpyAsync(d_out, tmp, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

// Exact order statistics x_(k_lo), x_(k_hi) (0-based, k_lo <= k_hi) into d_out[0], d_out[1].
// First level: linear bins across mean +- 1.0625 sigma (the median is within one
// sigma of the mean); the bins holding both ranks are gathered and sorted.
template <typename T>
void select_ranks(stp_ctx* ctx, const T* x, size_t n, double mean, double sigma, bool stats_valid, size_t k_lo, size_t k_hi,
    double* d_out)
{
    if (stats_valid && sigma > 0.0 && std::isfinite(sigma) && std::isfinite(mean)) {
        const unsigned nb = hist_blocks(n);
        uint32_t* bh = wsp<uint32_t>(ctx, "sf_bh", size_t(nb) * (LIN_BINS + 2));
        unsigned long long* hist = wsp<unsigned long long>(ctx, "sf_hist", LIN_BINS + 2);
        unsigned long long* cand = wsp<unsigned long long>(ctx, "sf_cand", SORT_CAP);
        unsigned* counter = wsp<unsigned>(ctx, "sf_counter", 1);
        const double lo = mean - 1.0625 * sigma, hi = mean + 1.0625 * sigma;
        const double scale = double(LIN_BINS) / (hi - lo);
        launch(ctx, "sf_lin_hist", sf_lin_hist<T>, dim3(nb), dim3(SWG), 0, x, n, lo, scale, bh);
        launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(LIN_BINS + 2, 256)), dim3(256), 0, (const uint32_t*)bh,
            int(nb), LIN_BINS + 2, hist);
        std::vector<unsigned long long> h(LIN_BINS + 2);
        d2h_sync(ctx, h.data(), hist, sizeof(unsigned long long) * h.size());
        size_t cum = 0, before_lo = 0;
        int blo = -1, bhi = -1;
        for (int b = 0; b < LIN_BINS + 2; b++) {
            if (blo < 0 && cum + h[b] > k_lo) {
                blo = b;
                before_lo = cum;
            }
            if (bhi < 0 && cum + h[b] > k_hi) {
                bhi = b;
                break;
            }
            cum += h[b];
        }
        size_t cnt = 0;
        if (blo >= 0 && bhi >= blo)
            for (int b = blo; b <= bhi; b++)
                cnt += h[b];
        if (blo > 0 && bhi <= LIN_BINS && bhi >= blo && cnt <= size_t(SORT_CAP)) {
            GatherSpec gs{};
            gs.mode = 0;
            gs.lo = lo;
            gs.scale = scale;
            gs.b_lo = blo;
            gs.b_hi = bhi;
            fill_u32(ctx, counter, 0u, 1);
            launch(ctx, "sf_gather", sf_gather<T>, dim3(nb), dim3(SWG), 0, x, n, gs, cand, counter, unsigned(SORT_CAP));
            launch(ctx, "sf_sort_select", sf_sort_select, dim3(1), dim3(1024), 0, (const unsigned long long*)cand,
                (const unsigned*)counter, int(k_lo - before_lo), int(k_hi - before_lo), d_out);
            return;
        }
    }
    radix_select<T>(ctx, x, n, k_lo, d_out);
    radix_select<T>(ctx, x, n, k_hi, d_out + 1);
}

void moments_fit(stp_island& isl, double x_bar, double y_bar, double xx_bar, double yy_bar, double xy_bar, double rms_est,
    double analysis_n_sigma)
{
    // IslandParams::estimate_moments_fit (sourcefind.cpp:691-717) and
    // Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57)
    const double working1 = (xx_bar + yy_bar) / 2.0;
    const double working2 = std::sqrt(((xx_bar - yy_bar) * (xx_bar - yy_bar) / 4) + xy_bar * xy_bar);
    const double major_sq = working1 + working2, minor_sq = working1 - working2;
    const double pixel_threshold = analysis_n_sigma * rms_est;
    const double cutoff_ratio = isl.sign * isl.extremum_val / pixel_threshold;
    const double axes_scale = 1.0 - std::log(cutoff_ratio) / (cutoff_ratio - 1.0);
    double semimajor = std::sqrt(major_sq / axes_scale), semiminor = std::sqrt(minor_sq / axes_scale);
    double theta = 0.5 * std::atan(2.0 * xy_bar / (xx_bar - yy_bar));
    if (theta * xy_bar < 0.0)
        theta += M_PI / 2.0;
    const double half_pi = M_PI / 2.0;
    semimajor = std::fabs(semimajor);
    semiminor = std::fabs(semiminor);
    if (semimajor < semiminor) {
        std::swap(semimajor, semiminor);
        theta += half_pi;
    }
    double mt = std::fmod(theta, M_PI);
    if (mt <= -half_pi)
        mt += M_PI;
    else if (mt > half_pi)
        mt -= M_PI;
    isl.moments_fit = stp_gaussian2d{ isl.extremum_val, x_bar, y_bar, semimajor, semiminor, mt };
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
    SFState hs{};
    const bool need_rms = !(std::fabs(p.rms_est) > 0.0);
    const int mm = p.median_method;
    bool stats_valid = false;
    auto compute_mean_sigma = [&]() {
        launch(ctx, "sf_sum", sf_sum_kernel<T>, dim3(unsigned(nblk)), dim3(SWG), 0, x, n, part);
        canon_reduce_pairs(ctx, part, nblk, scr, sums);
        launch(ctx, "sf_stats", sf_mean_sigma_kernel, dim3(1), dim3(1), 0, (const double*)sums, (unsigned long long)n, st);
        stats_valid = true;
    };
    auto exact_median = [&]() {
        if (stats_valid)
            d2h_sync(ctx, &hs, st, sizeof(SFState));
        const size_t k_hi = n / 2, k_lo = (n & 1) ? n / 2 : n / 2 - 1;
        select_ranks<T>(ctx, x, n, hs.mean, hs.sigma, stats_valid, k_lo, k_hi, sel);
        launch(ctx, "sf_median", sf_set_median, dim3(1), dim3(1), 0, st, (const double*)sel, int(n & 1));
    };
    fill_f64(ctx, (double*)st, 0.0, sizeof(SFState) / sizeof(double));
    // median (sourcefind.cpp:182-197)
    if (mm == STP_MEDIAN_BINAPPROX) {
        compute_mean_sigma();
        d2h_sync(ctx, &hs, st, sizeof(SFState));
        if ((hs.sigma * 2) < 0.0001) {
            exact_median();
        } else {
            const unsigned nb = hist_blocks(n);
            uint32_t* bh = wsp<uint32_t>(ctx, "sf_babh", size_t(nb) * (BA_BINS + 1));
            unsigned long long* hist = wsp<unsigned long long>(ctx, "sf_bahist", BA_BINS + 1);
            int* found = wsp<int>(ctx, "sf_found", 1);
            launch(ctx, "sf_binapprox_hist", sf_binapprox_hist<T>, dim3(nb), dim3(SWG), 0, x, n, (const SFState*)st, bh);
            launch(ctx, "sf_hist_reduce", sf_hist_reduce, dim3(div_up(BA_BINS + 1, 256)), dim3(256), 0,
                (const uint32_t*)bh, int(nb), BA_BINS + 1, hist);
            launch(ctx, "sf_binapprox_pick", sf_binapprox_pick, dim3(1), dim3(1), 0, (const unsigned long long*)hist,
                (unsigned long long)n, st, found);
            int hf = 0;
            d2h_sync(ctx, &hf, found, sizeof(int));
            if (!hf)
                exact_median();
        }
    } else if (mm == STP_MEDIAN_BINMEDIAN) {
        compute_mean_sigma();
        exact_median();
    } else if (mm == STP_MEDIAN_NTHELEMENT) {
        // mean and sigma are needed by estimate_rms anyway (sourcefind.cpp:30-36);
        // they also seed the linear-bin first level of the selection
        if (need_rms)
            compute_mean_sigma();
        exact_median();
    }
    // ZEROMEDIAN: the state's median stays 0
    // RMS (sourcefind.cpp:206)
    if (need_rms) {
        if (!stats_valid)
            compute_mean_sigma();
        const int iters = std::max(0, p.sigma_clip_iters);
        unsigned long long* removed = wsp<unsigned long long>(ctx, "sf_removed", std::max(1, iters));
        double* cpart = wsp<double>(ctx, "sf_cpart", 2 * nblk + 2);
        double* csums = wsp<double>(ctx, "sf_csums", 2);
        launch(ctx, "sf_clip_init", sf_clip_init, dim3(1), dim3(1), 0, st, (unsigned long long)n, iters, 3.0);
        if (iters > 0) {
            fill_f64(ctx, (double*)removed, 0.0, size_t(iters));
            for (int it = 0; it < iters; it++) {
                launch(ctx, "sf_clip", sf_clip_kernel<T>, dim3(unsigned(nblk)), dim3(SWG), 0, x, n, (const SFState*)st, cpart, removed);
                canon_reduce_pairs(ctx, cpart, nblk, scr, csums);
                launch(ctx, "sf_clip_update", sf_clip_update, dim3(1), dim3(1), 0, st, (const double*)csums,
                    (const unsigned long long*)removed);
            }
        }
    }
    launch(ctx, "sf_thresholds", sf_thresholds, dim3(1), dim3(1), 0, st, int(need_rms), p.rms_est, p.detection_n_sigma,
        p.analysis_n_sigma);

    // connected components (ccl.h), ids in the reference's scan order
    uint8_t* cls = wsp<uint8_t>(ctx, "ccl_cls", n);
    uint32_t* P = wsp<uint32_t>(ctx, "ccl_parent", n);
    int32_t* ID = wsp<int32_t>(ctx, "ccl_id", n);
    int32_t* lab = d_lab_out ? d_lab_out : wsp<int32_t>(ctx, "ccl_label", n);
    launch(ctx, "ccl_classify", ccl_classify<T>, dim3(div_up(n, SWG)), dim3(SWG), 0, x, rows, cols, (const SFState*)st,
        int(p.find_negative_sources != 0), cls, P);
    launch(ctx, "ccl_merge", ccl_merge, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, P, rows, cols,
        int(p.ccl_4connectivity != 0));
    launch(ctx, "ccl_flatten", ccl_flatten, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, P, n);
    uint32_t* cpos = wsp<uint32_t>(ctx, "ccl_cpos", nblk);
    uint32_t* cneg = wsp<uint32_t>(ctx, "ccl_cneg", nblk);
    uint32_t* opos = wsp<uint32_t>(ctx, "ccl_opos", nblk + 1);
    uint32_t* oneg = wsp<uint32_t>(ctx, "ccl_oneg", nblk + 1);
    uint32_t* cur = wsp<uint32_t>(ctx, "ccl_cur", nblk);
    launch(ctx, "ccl_root_count", ccl_root_count, dim3(unsigned(nblk)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P,
        n, cpos, cneg);
    exclusive_scan_u32(ctx, cpos, int(nblk), opos, cur);
    exclusive_scan_u32(ctx, cneg, int(nblk), oneg, cur);
    launch(ctx, "ccl_root_ids", ccl_root_ids, dim3(unsigned(nblk)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P, n,
        (const uint32_t*)opos, (const uint32_t*)oneg, ID);
    launch(ctx, "ccl_label_out", ccl_label_out, dim3(div_up(n, SWG)), dim3(SWG), 0, (const uint8_t*)cls, (const uint32_t*)P,
        (const int32_t*)ID, rows, cols, lab);
    uint32_t kpk[2] = { 0, 0 };
    HIP_CHECK(hipMemcpyAsync(&kpk[0], opos + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(&kpk[1], oneg + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const uint32_t kp = kpk[0], kn = kpk[1];

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

    // results to the host
    std::vector<unsigned long long> hep(kp), hen(kn), hip_(kp), hin(kn);
    std::vector<int> hcp(kp), hcn(kn);
    std::vector<double> hmp(6 * size_t(kp)), hmn(6 * size_t(kn));
    auto cp_d2h = [&](void* dst, const void* src, size_t bytes) {
        if (bytes)
            HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    cp_d2h(hep.data(), ep, 8 * size_t(kp));
    cp_d2h(hen.data(), en, 8 * size_t(kn));
    cp_d2h(hip_.data(), ip, 8 * size_t(kp));
    cp_d2h(hin.data(), in_, 8 * size_t(kn));
    cp_d2h(hcp.data(), cp, 4 * size_t(kp));
    cp_d2h(hcn.data(), cn, 4 * size_t(kn));
    cp_d2h(hmp.data(), mp, 48 * size_t(kp));
    cp_d2h(hmn.data(), mn, 48 * size_t(kn));
    cp_d2h(&hs, st, sizeof(SFState));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));

    auto key_inv = [](unsigned long long k) {
        const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        double d;
        std::memcpy(&d, &b, 8);
        return d;
    };
    size_t found = 0;
    uint32_t valid = 0;
    const int h_shift = cols / 2, v_shift = rows / 2;
    auto emit = [&](uint32_t count, const std::vector<unsigned long long>& e, const std::vector<unsigned long long>& idx,
                    const std::vector<int>& c, const std::vector<double>& m, int sign) {
        for (uint32_t l = 0; l < count; l++) {
            const double ext = key_inv(e[l]);
            const bool ok = sign > 0 ? (ext > hs.d_pos) : (ext < hs.d_neg);
            if (!ok)
                continue;
            valid++;
            if (c[l] < p.source_min_area)
                continue;
            if (found < cap) {
                stp_island isl{};
                const size_t li = size_t(idx[l]);
                const int r = int(li % size_t(rows)), cc = int(li / size_t(rows));
                isl.label_idx = sign > 0 ? int32_t(l + 1) : -int32_t(l + 1);
                isl.sign = sign;
                isl.extremum_y_idx = r < v_shift ? r + v_shift : r - v_shift;
                isl.extremum_x_idx = cc < h_shift ? cc + h_shift : cc - h_shift;
                isl.num_samples = c[l];
                isl.extremum_val = ext;
                const double* mm = &m[6 * size_t(l)];
                const double xb = mm[0] / mm[5], yb = mm[1] / mm[5];
                moments_fit(isl, xb, yb, mm[2] / mm[5] - xb * xb, mm[3] / mm[5] - yb * yb, mm[4] / mm[5] - xb * yb, hs.rms,
                    p.analysis_n_sigma);
                islands[found] = isl;
            }
            found++;
        }
    };
    emit(kp, hep, hip_, hcp, hmp, +1);
    emit(kn, hen, hin, hcn, hmn, -1);
    *n_islands = found;
    if (stats) {
        stats->bg_level = hs.median;
        stats->rms_est = hs.rms;
        stats->mean = stats_valid ? hs.mean : 0.0;
        stats->sigma = stats_valid ? hs.sigma : 0.0;
        stats->num_labels_pos = kp;
        stats->num_labels_neg = kn;
        stats->num_valid_labels = valid;
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
