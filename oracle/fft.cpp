/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * FFTs with the semantics of the reference's FFTW wrappers
 * (src/stp/common/fft.cpp:45-135 c2r, :227-323 c2c, :325-412 R2HC).
 * Own iterative radix-2 Cooley-Tukey transforms (FFTW is not available here).
 */
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "oracle.h"

namespace oracle {

int default_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    return n ? int(n) : 1;
}

template <typename F>
static void parallel_for(size_t n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 2) {
        for (size_t i = 0; i < n; i++)
            f(i);
        return;
    }
    const size_t nt = std::min<size_t>(size_t(nthreads), n);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++) {
        th.emplace_back([=, &f]() {
            const size_t b = n * t / nt, e = n * (t + 1) / nt;
            for (size_t i = b; i < e; i++)
                f(i);
        });
    }
    for (auto& x : th)
        x.join();
}

// Twiddles exp(-2*pi*i*k/n) for k < n/2, computed once per n.
static const std::vector<cx>& twiddles(size_t n)
{
    static std::mutex mu;
    static std::map<size_t, std::vector<cx>> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(n);
    if (it != cache.end())
        return it->second;
    std::vector<cx> t(n / 2 + 1);
    for (size_t k = 0; k < t.size(); k++) {
        const double a = -2.0 * M_PI * double(k) / double(n);
        t[k] = cx(std::cos(a), std::sin(a));
    }
    return cache.emplace(n, std::move(t)).first->second;
}

void fft1d(cx* a, size_t n, int sign)
{
    if (n < 2)
        return;
    if (n & (n - 1))
        throw std::invalid_argument("fft1d: length must be a power of two");
    // bit reversal permutation
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j)
            std::swap(a[i], a[j]);
    }
    const std::vector<cx>& tw = twiddles(n);
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2, step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < half; k++) {
                cx w = tw[k * step];
                if (sign > 0)
                    w = std::conj(w);
                const cx u = a[i + k];
                const cx v = a[i + k + half] * w;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

void fft_c2r_2d(const cx* in, int in_rows, int cols, double* out, int nthreads)
{
    // n_rows of the real output (fft.cpp:47): even input rows would be 2*in_rows;
    // the imager always passes in_rows = n/2+1 (odd for n >= 2).
    if (in_rows < 2 || (in_rows % 2) == 0)
        throw std::invalid_argument("fft_c2r_2d: expects n/2+1 input rows");
    const int n = (in_rows - 1) * 2;
    std::vector<cx> h(size_t(in_rows) * cols);
    // c2c (sign +1) along u for every v row
    parallel_for(size_t(in_rows), nthreads, [&](size_t v) {
        std::vector<cx> row(cols);
        for (int u = 0; u < cols; u++)
            row[u] = in[v + size_t(u) * in_rows];
        fft1d(row.data(), cols, +1);
        for (int x = 0; x < cols; x++)
            h[v + size_t(x) * in_rows] = row[x];
    });
    // c2r along v for every output column; imaginary parts of v=0 and v=n/2 ignored
    parallel_for(size_t(cols), nthreads, [&](size_t x) {
        std::vector<cx> col(n);
        const cx* hc = &h[size_t(x) * in_rows];
        col[0] = cx(hc[0].real(), 0.0);
        col[n / 2] = cx(hc[n / 2].real(), 0.0);
        for (int v = 1; v < n / 2; v++) {
            col[v] = hc[v];
            col[n - v] = std::conj(hc[v]);
        }
        fft1d(col.data(), n, +1);
        for (int y = 0; y < n; y++)
            out[y + size_t(x) * n] = col[y].real();
    });
}

void fft_c2c_2d(const cx* in, int rows, int cols, bool forward, cx* out, int nthreads)
{
    const int sign = forward ? -1 : +1;
    std::vector<cx> tmp(in, in + size_t(rows) * cols);
    parallel_for(size_t(cols), nthreads, [&](size_t c) { fft1d(&tmp[c * rows], rows, sign); });
    parallel_for(size_t(rows), nthreads, [&](size_t r) {
        std::vector<cx> row(cols);
        for (int c = 0; c < cols; c++)
            row[c] = tmp[r + size_t(c) * rows];
        fft1d(row.data(), cols, sign);
        for (int c = 0; c < cols; c++)
            out[r + size_t(c) * rows] = row[c];
    });
}

void fft_r2hc_1d(const double* in, size_t n, double* out)
{
    // FFTW_R2HC: r0, r1, ..., r_{n/2}, i_{(n+1)/2-1}, ..., i_1 of the forward DFT
    std::vector<cx> a(n);
    for (size_t i = 0; i < n; i++)
        a[i] = cx(in[i], 0.0);
    if (n & (n - 1)) {
        // direct DFT for non power-of-two sizes
        std::vector<cx> b(n);
        for (size_t k = 0; k < n; k++) {
            cx s(0.0, 0.0);
            for (size_t j = 0; j < n; j++) {
                const double ang = -2.0 * M_PI * double((k * j) % n) / double(n);
                s += in[j] * cx(std::cos(ang), std::sin(ang));
            }
            b[k] = s;
        }
        a = b;
    } else {
        fft1d(a.data(), n, -1);
    }
    for (size_t k = 0; k <= n / 2; k++)
        out[k] = a[k].real();
    for (size_t k = 1; k < (n + 1) / 2; k++)
        out[n - k] = a[k].imag();
}

} // namespace oracle
