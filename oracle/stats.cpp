/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Image statistics: src/stp/common/matrix_math.cpp:20-499 and the sigma-clipped
 * RMS of src/stp/sourcefind/sourcefind.cpp:22-154.
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>
#include <vector>

#include "oracle.h"

namespace oracle {

/*
 * Canonical summation order.  The reference sums with TBB parallel_reduce /
 * combinable, whose association order depends on the thread schedule
 * (matrix_math.cpp:467-494, sourcefind.cpp:58-144).  This build fixes one
 * order, shared bit-for-bit with the HIP reduction kernels:
 *   - the input is cut into blocks of 4096 consecutive elements;
 *   - inside a block, lane t (0..255) adds elements t, t+256, ..., t+15*256
 *     sequentially starting from 0.0;
 *   - the 256 lane sums are combined by the tree p[t] = p[t] + p[t+off] for
 *     off = 128, 64, ..., 1 (t < off); the block result is p[0];
 *   - the block results are reduced again with the same scheme until one
 *     value is left.
 */
static const size_t CANON_BLOCK = 4096, CANON_LANES = 256, CANON_PER_LANE = 16;

struct Pair {
    double a, b;
};

static void block_tree(Pair* p)
{
    for (size_t off = CANON_LANES / 2; off >= 1; off /= 2)
        for (size_t t = 0; t < off; t++) {
            p[t].a = p[t].a + p[t + off].a;
            p[t].b = p[t].b + p[t + off].b;
        }
}

// one level over pairs (already summed terms)
static std::vector<Pair> reduce_level(const std::vector<Pair>& in)
{
    const size_t n = in.size(), nb = (n + CANON_BLOCK - 1) / CANON_BLOCK;
    std::vector<Pair> out(nb);
    Pair p[CANON_LANES];
    for (size_t b = 0; b < nb; b++) {
        for (size_t t = 0; t < CANON_LANES; t++) {
            double s1 = 0.0, s2 = 0.0;
            for (size_t j = 0; j < CANON_PER_LANE; j++) {
                const size_t i = b * CANON_BLOCK + j * CANON_LANES + t;
                if (i < n) {
                    s1 += in[i].a;
                    s2 += in[i].b;
                }
            }
            p[t] = { s1, s2 };
        }
        block_tree(p);
        out[b] = p[0];
    }
    return out;
}

static Pair finish(std::vector<Pair> v)
{
    if (v.empty())
        return { 0.0, 0.0 };
    while (v.size() > 1)
        v = reduce_level(v);
    return v[0];
}

// Level 0 over raw terms produced by term(i) -> (a_i, b_i).
template <typename F>
static Pair canonical_terms(size_t n, F term)
{
    const size_t nb = (n + CANON_BLOCK - 1) / CANON_BLOCK;
    std::vector<Pair> out(nb);
    Pair p[CANON_LANES];
    for (size_t b = 0; b < nb; b++) {
        for (size_t t = 0; t < CANON_LANES; t++) {
            double s1 = 0.0, s2 = 0.0;
            for (size_t j = 0; j < CANON_PER_LANE; j++) {
                const size_t i = b * CANON_BLOCK + j * CANON_LANES + t;
                if (i < n) {
                    const Pair q = term(i);
                    s1 += q.a;
                    s2 += q.b;
                }
            }
            p[t] = { s1, s2 };
        }
        block_tree(p);
        out[b] = p[0];
    }
    return finish(std::move(out));
}

void canonical_pair_sum(const double* x, size_t n, double* s1, double* s2)
{
    const Pair r = canonical_terms(n, [&](size_t i) {
        const double v = x[i];
        return Pair{ v, v * v };
    });
    *s1 = r.a;
    *s2 = r.b;
}

// mat_mean_and_stddev (matrix_math.cpp:457-499)
Stats mat_mean_and_stddev(const double* x, size_t n)
{
    Stats s;
    if (n < 1)
        return s;
    double t1, t2;
    canonical_pair_sum(x, n, &t1, &t2);
    const double mean = t1 / double(n);
    const double sigma = std::sqrt(t2 / double(n) - (mean * mean));
    s.mean = mean;
    s.sigma = sigma;
    s.mean_valid = s.sigma_valid = true;
    return s;
}

// mat_median_exact (matrix_math.cpp:20-48)
double mat_median_exact(const double* x, size_t n)
{
    if (n == 0)
        return 0.0;
    std::vector<double> a(x, x + n);
    const size_t k = n / 2;
    std::nth_element(a.begin(), a.begin() + k, a.end());
    if (n & 1)
        return a[k];
    const double val1 = a[k];
    const double val2 = *std::max_element(a.begin(), a.begin() + k);
    return (val1 + val2) / 2;
}

// mat_binmedian (matrix_math.cpp:50-332): successive binning narrows onto the
// bin holding the middle order statistics and finishes with nth_element over
// that bin, i.e. the exact median (x_(n/2-1) + x_(n/2))/2 (even n) or x_(n/2)
// (odd n, 0-based), plus the one-pass mean and sigma.
Stats mat_binmedian(const double* x, size_t n)
{
    Stats s = mat_mean_and_stddev(x, n);
    s.median = mat_median_exact(x, n);
    s.median_valid = true;
    return s;
}

/*
 * Bin index int((x - leftend) * scalefactor) of matrix_math.cpp:370-392.  The
 * reference converts a double to int with truncation; for values outside int
 * range the x86 conversion instruction returns INT_MIN (counted as below the
 * range).  That behaviour is made explicit here so the GPU kernels can match it.
 */
static inline int bin_index(double x, double leftend, double scalefactor)
{
    const double t = (x - leftend) * scalefactor;
    if (!(t > -2147483649.0 && t < 2147483648.0))
        return INT_MIN;
    return int(t);
}

// mat_median_binapprox (matrix_math.cpp:334-455)
Stats mat_median_binapprox(const double* x, size_t n)
{
    const int N = 1000;
    Stats s = mat_mean_and_stddev(x, n);
    const double mean = s.mean, sigma = s.sigma;
    s.median_valid = true;
    if ((sigma * 2) < 0.0001) {
        s.median = mat_median_exact(x, n);
        return s;
    }
    const double scalefactor = double(N) / (2 * sigma);
    const double leftend = mean - sigma;
    std::vector<size_t> bins(N + 1, 0);
    size_t bottom = 0;
    for (size_t i = 0; i < n; i++) {
        const int b = bin_index(x[i], leftend, scalefactor);
        if (b < 0)
            bottom++;
        else if (b <= N)
            bins[b]++;
    }
    bool ok = false;
    double median = 0.0;
    if (n & 1) {
        const size_t k = (n + 1) / 2;
        size_t count = bottom;
        for (unsigned i = 0; i < unsigned(N + 1); i++) {
            count += bins[i];
            if (count >= k) {
                median = double(i + 0.5) / scalefactor + leftend;
                ok = true;
                break;
            }
        }
    } else {
        const size_t k = n / 2;
        size_t count = bottom;
        for (unsigned i = 0; i < unsigned(N + 1); i++) {
            count += bins[i];
            if (count >= k) {
                unsigned j = i;
                while (count == k) {
                    j++;
                    count += bins[j];
                    if (j == unsigned(N))
                        break;
                }
                median = double(i + j + 1) / (2 * scalefactor) + leftend;
                ok = true;
                break;
            }
        }
    }
    s.median = ok ? median : mat_median_exact(x, n);
    return s;
}

// estimate_rms (sourcefind.cpp:22-154).  The integer compare of FloatTwiddler
// bit patterns (:84-90) equals the float compare |val| > bound for finite data.
double estimate_rms(const double* x, size_t n, double num_sigma, unsigned iters, Stats st)
{
    if (!st.median_valid) {
        st = mat_binmedian(x, n);
    } else if (!st.mean_valid || !st.sigma_valid) {
        const Stats t = mat_mean_and_stddev(x, n);
        st.mean = t.mean;
        st.sigma = t.sigma;
        st.mean_valid = st.sigma_valid = true;
    }
    const double median = st.median;
    const double mean = st.mean - median;
    double sigma = st.sigma;
    double d1 = double(mean) * double(n);
    double d2 = (double(sigma) * double(sigma) + double(mean) * double(mean)) * double(n);
    size_t valid_n = n;
    double prev_upper = std::numeric_limits<double>::max();
    for (unsigned it = 0; it < iters; it++) {
        const double upper = num_sigma * sigma;
        size_t removed = 0;
        for (size_t i = 0; i < n; i++) {
            const double a = std::fabs(x[i] - median);
            if (a > upper && !(a > prev_upper))
                removed++;
        }
        if (removed == 0)
            break;
        const Pair r = canonical_terms(n, [&](size_t i) {
            const double v = x[i] - median;
            const double a = std::fabs(v);
            if (a > upper && !(a > prev_upper))
                return Pair{ v, v * v };
            return Pair{ 0.0, 0.0 };
        });
        valid_n -= removed;
        d1 -= r.a;
        d2 -= r.b;
        sigma = std::sqrt(d2 / double(valid_n) - (d1 / double(valid_n)) * (d1 / double(valid_n)));
        prev_upper = upper;
    }
    return sigma;
}

} // namespace oracle
