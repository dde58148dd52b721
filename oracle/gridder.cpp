/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Convolutional gridder: src/stp/gridder/gridder.h:73-105,240-997,
 * src/stp/gridder/gridder.cpp:11-146,202-241,
 * W-projection kernels: src/stp/gridder/aw_projection.cpp:19-111,173-210,366-474.
 */
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "oracle.h"

namespace oracle {

// populate_kernel_cache (gridder.h:73-105): cache[a * cs + b] is the
// (2s+1)^2 column-major matrix k_a[row] * k_b[col] of the 1-D kernels at
// oversampled offsets (a - os/2)/os (rows) and (b - os/2)/os (columns).
std::vector<std::vector<cx>> populate_kernel_cache(const KernelFunc& kf, unsigned support, unsigned oversampling)
{
    const unsigned op = oversampling / 2, cs = op * 2 + 1;
    std::vector<std::vector<double>> k1d(cs);
    for (unsigned i = 0; i < cs; i++) {
        const double off = (double(i) - double(op)) / double(oversampling);
        k1d[i] = make_1d_kernel(kf, int(support), off, 1, false, true);
    }
    const size_t ksz = k1d[0].size();
    std::vector<std::vector<cx>> cache(size_t(cs) * cs, std::vector<cx>(ksz * ksz));
    for (unsigned a = 0; a < cs; a++)
        for (unsigned b = 0; b < cs; b++)
            for (size_t c = 0; c < ksz; c++)
                for (size_t r = 0; r < ksz; r++)
                    cache[a * cs + b][r + c * ksz] = cx(k1d[a][r] * k1d[b][c], 0.0);
    return cache;
}

// convert_to_halfplane_visibilities (gridder.cpp:11-66)
void convert_to_halfplane_visibilities(std::vector<double>& uv, size_t n, std::vector<double>* w, std::vector<cx>& vis, int support, std::vector<uint32_t>& good)
{
    for (size_t i = 0; i < n; i++) {
        const double v = uv[n + i];
        if (v < 0.0) {
            if (v > -(support + 1))
                good[i] = 2;
            uv[i] *= -1;
            uv[n + i] *= -1;
            if (w)
                (*w)[i] *= -1;
            vis[i] = std::conj(vis[i]);
        } else {
            if (v < (support + 1))
                good[i] = 2;
        }
    }
}

// bounds_check_kernel_centre_locations (gridder.cpp:202-217); kc is n x 2 column-major.
void bounds_check_kernel_centre_locations(std::vector<uint32_t>& good, const std::vector<int>& kc, size_t n, int image_size, int support)
{
    for (size_t i = 0; i < n; i++) {
        const int x = kc[i], y = kc[n + i];
        if ((x - support) <= 0 || (y - support) <= 0 || (x + support) >= image_size || (y + support) >= image_size)
            good[i] = 0;
    }
}

// calculate_oversampled_kernel_indices (gridder.cpp:219-241)
void calculate_oversampled_kernel_indices(const double* frac, size_t n, unsigned oversampling, int* out)
{
    const int range_max = int(oversampling / 2);
    const int range_min = -range_max;
    for (size_t i = 0; i < n; i++) {
        int val = int(std::rint(frac[i] * oversampling));
        if (val > range_max)
            val = range_max;
        else if (val < range_min)
            val = range_min;
        out[i] = val;
    }
}

// average_w_planes (gridder.cpp:68-146)
void average_w_planes(const std::vector<double>& w, const std::vector<uint32_t>& good, unsigned num_wplanes,
    std::vector<double>& w_avg, std::vector<size_t>& first_idx, bool median)
{
    const size_t num_elems = good.size();
    size_t num_gvis = 0;
    for (auto g : good)
        if (g != 0)
            num_gvis++;
    w_avg.assign(num_wplanes, 0.0);
    first_idx.assign(num_wplanes, 0);
    size_t begin = 0;
    for (unsigned idx = 0; idx < num_wplanes; idx++) {
        const size_t plane_size = size_t(std::ceil(double(num_gvis - begin) / (num_wplanes - idx)));
        size_t k, counter = 0;
        if (median) {
            size_t idx_mid = 0, middle = plane_size / 2;
            if ((begin + plane_size) > num_gvis)
                middle = (num_gvis - begin) / 2;
            for (k = begin; (counter < plane_size) && (k < num_elems); k++) {
                if (good[k] != 0) {
                    if (counter == middle)
                        idx_mid = k;
                    counter++;
                }
            }
            if (counter > 0)
                w_avg[idx] = (plane_size % 2 == 0) ? (w[idx_mid - 1] + w[idx_mid]) / 2.0 : w[idx_mid];
            else
                w_avg[idx] = 0.0;
            first_idx[idx] = begin;
            begin = k;
        } else {
            double sum = 0;
            for (k = begin; (counter < plane_size) && (k < num_elems); k++) {
                if (good[k] != 0) {
                    sum += w[k];
                    counter++;
                }
            }
            first_idx[idx] = begin;
            begin = k;
            w_avg[idx] = counter > 0 ? sum / counter : 0.0;
        }
    }
}

// WideFieldImaging::combine_2d_kernel_value (aw_projection.cpp:455-474)
static cx combine_2d_kernel_value(const std::vector<double>& aa, size_t px, size_t py, double w_value, double scaled_cell, unsigned ctr)
{
    const double dx = double(px) * scaled_cell;
    const double dy = double(py) * scaled_cell;
    const double r2 = dx * dx + dy * dy;
    if (r2 >= 1.0)
        return cx(aa[ctr + px] * aa[ctr + py], 0.0);
    const double nn = std::sqrt(1.0 - r2);
    const double ph = (-2.0 * M_PI) * w_value * (nn - 1.0);
    cx v = cx(std::cos(ph), std::sin(ph)) / nn;
    v = v * aa[ctr + px];
    v = v * aa[ctr + py];
    return v;
}

std::vector<std::vector<cx>> wproj_kernel_cache(double w_value, const std::vector<double>& aa, unsigned workarea_size,
    double cell_rad, unsigned oversampling, double scaling_factor, unsigned support, int nthreads, double trunc_perc,
    unsigned* trunc_support)
{
    // generate_image_domain_convolution_kernel (aw_projection.cpp:50-111)
    const size_t arr = size_t(workarea_size) * oversampling;
    const unsigned half = workarea_size / 2;
    const double scaled_cell = cell_rad * scaling_factor;
    std::vector<cx> comb(arr * arr, cx(0.0, 0.0));
    auto at = [&](size_t i, size_t j) -> cx& { return comb[i + j * arr]; };
    for (size_t i = 1; i < half; i++) {
        const cx v = combine_2d_kernel_value(aa, i, 0, w_value, scaled_cell, half);
        at(i, 0) = v;
        at(arr - i, 0) = v;
        at(0, i) = v;
        at(0, arr - i) = v;
    }
    at(0, 0) = cx(aa[half] * aa[half], 0.0);
    for (size_t j = 1; j < half; j++)
        for (size_t i = 1; i < half; i++) {
            const cx v = combine_2d_kernel_value(aa, i, j, w_value, scaled_cell, half);
            at(i, j) = v;
            at(arr - i, j) = v;
            at(i, arr - j) = v;
            at(arr - i, arr - j) = v;
        }
    // forward 2-D c2c (aw_projection.cpp:194)
    std::vector<cx> conv(arr * arr);
    fft_c2c_2d(comb.data(), int(arr), int(arr), true, conv.data(), nthreads);
    // truncation at kernel_trunc_perc % of the peak (aw_projection.cpp:176-179,197-209):
    // the largest support whose sample conv(s * os, 0) still exceeds that fraction
    if (trunc_perc > 0.0) {
        const double min_value = std::abs(conv[0]) * (trunc_perc / 100.0);
        size_t i = std::min(size_t(support) * oversampling, arr >> 1);
        while (i > 0) {
            if (std::abs(conv[i]) > min_value)
                break;
            i -= oversampling;
        }
        support = std::min(unsigned(i / oversampling), support);
        if (support == 0)
            throw std::invalid_argument("W-kernel truncated to support 0 (gridder.h:584 asserts support > 0)");
    }
    if (trunc_support)
        *trunc_support = support;
    // generate_kernel_cache, non-Hankel branch (aw_projection.cpp:415-443)
    const size_t ctr_idx = arr / 2, op = oversampling / 2, cache_size = 2 * op + 1;
    const size_t tcs = 2 * size_t(support) + 1;
    const size_t oversamp_conv = ctr_idx - size_t(support) * oversampling + op;
    std::vector<std::vector<cx>> cache(cache_size * cache_size, std::vector<cx>(tcs * tcs));
    for (size_t x = 0; x < cache_size; x++) {
        for (size_t y = 0; y < cache_size; y++) {
            std::vector<cx>& tmp = cache[y * cache_size + x];
            double ksum = 0;
            for (size_t ty = 0; ty < tcs; ty++) {
                const size_t cxi = (oversamp_conv - x + ty * oversampling + ctr_idx) % arr;
                for (size_t tx = 0; tx < tcs; tx++) {
                    const size_t cyi = (oversamp_conv - y + tx * oversampling + ctr_idx) % arr;
                    const cx v = conv[cxi + cyi * arr];
                    tmp[tx + ty * tcs] = v;
                    ksum += v.real();
                }
            }
            ksum = 1 / ksum;
            for (auto& v : tmp)
                v = cx(v.real() * ksum, v.imag() * ksum);
        }
    }
    return cache;
}

// ---------------- Hankel W-kernels (wp.hankel_opt) ----------------

// besselj1 (aw_projection.cpp:490-508): J1 by the rational approximation of
// Numerical Recipes in C, 2nd ed., section 6.5 (bessj1), whose constants the
// reference uses; |error| < 1e-7.
double besselj1(double x)
{
    if (x < 8.0) {
        const double y = x * x;
        const double p = x * (72362614232.0 + y * (-7895059235.0 + y * (242396853.1 + y * (-2972611.439 + y * (15704.48260 + y * (-30.16036606))))));
        const double q = 144725228442.0 + y * (2300535178.0 + y * (18583304.74 + y * (99447.43394 + y * (376.9991397 + y))));
        return p / q;
    }
    const double z = 8.0 / x, y = z * z, xx = x - 2.356194491;
    const double p = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4 + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
    const double q = 0.04687499995 + y * (-0.2002690873e-3 + y * (0.8449199096e-5 + y * (-0.88228987e-6 + y * 0.105787412e-6)));
    return std::sqrt(0.636619772 / x) * (std::cos(xx) * p - z * std::sin(xx) * q);
}

// dht (aw_projection.cpp:530-596): the n x n discrete Hankel transform, column-major
// D[j + i n] = DHT(j, i): row j of column i is the integral of 2 pi r J0(k_i r) over
// the ring [rn_{j-1}, rn_j] (rn_j = j + 1/2, the last one n - 1), from the primitive
// (2 pi / k) r J1(k r); column 0 (k = 0) holds the ring areas.
std::vector<double> dht_matrix(size_t n)
{
    std::vector<double> k(n), rn(n), kn(n), D(n * n);
    const double kf = M_PI / double(n);
    kn[0] = 0.0;
    k[0] = 0.0;
    rn[0] = 0.5;
    for (size_t i = 1; i < n; i++) {
        k[i] = kf * double(i);
        rn[i] = double(i) + 0.5;
        kn[i] = (2.0 * M_PI) / k[i];
    }
    rn[n - 1] = double(n - 1);
    D[0] = M_PI * rn[0] * rn[0];
    double prev = D[0];
    for (size_t j = 1; j < n; j++) {
        const double cur = M_PI * rn[j] * rn[j];
        D[j] = cur - prev;
        prev = cur;
    }
    for (size_t i = 1; i < n; i++) {
        double* col = D.data() + i * n;
        col[n - 1] = kn[i] * rn[n - 1] * besselj1(k[i] * rn[n - 1]);
        for (size_t j = n - 1; j > 0; j--) {
            col[j - 1] = kn[i] * rn[j - 1] * besselj1(k[i] * rn[j - 1]);
            col[j] -= col[j - 1];
        }
    }
    return D;
}

// tk::spline on the knots x_i = i (spline.h, the reference's vendored cubic/linear
// spline): natural cubic (zero curvature at both ends: the tridiagonal system
// solved by forward elimination and back substitution) or linear; coefficients a,
// b, c with f(x) = ((a h + b) h + c) h + y_k, h = x - k.
void spline_coefs(const double* y, size_t n, bool cubic, double* a, double* b, double* c)
{
    if (!cubic) {
        for (size_t i = 0; i + 1 < n; i++) {
            c[i] = (y[i + 1] - y[i]) / 1.0;
            a[i] = b[i] = 0.0;
        }
        c[n - 1] = c[n - 2];
        a[n - 1] = b[n - 1] = 0.0;
        return;
    }
    // rows 1..n-2: (1/3) b_{i-1} + (2/3)(2) b_i + (1/3) b_{i+1} = (y_{i+1} - y_i) - (y_i - y_{i-1});
    // rows 0 and n-1: 2 b = 0
    const double lo = 1.0 / 3.0 * 1.0, di = 2.0 / 3.0 * 2.0, up = 1.0 / 3.0 * 1.0;
    std::vector<double> d(n), r(n);
    d[0] = 2.0;
    r[0] = 0.0;
    for (size_t i = 1; i < n; i++) {
        const bool last = i == n - 1;
        const double li = last ? 0.0 : lo, dii = last ? 2.0 : di;
        const double ri = last ? 0.0 : (y[i + 1] - y[i]) / 1.0 - (y[i] - y[i - 1]) / 1.0;
        const double upper_prev = (i == 1) ? 0.0 : up; // A(i-1, i)
        const double m = li / d[i - 1];
        d[i] = dii - m * upper_prev;
        r[i] = ri - m * r[i - 1];
    }
    b[n - 1] = r[n - 1] / d[n - 1];
    for (size_t i = n - 1; i-- > 0;) {
        const double upper = (i == 0) ? 0.0 : up;
        b[i] = (r[i] - upper * b[i + 1]) / d[i];
    }
    for (size_t i = 0; i + 1 < n; i++) {
        a[i] = 1.0 / 3.0 * (b[i + 1] - b[i]) / 1.0;
        c[i] = (y[i + 1] - y[i]) / 1.0 - 1.0 / 3.0 * (2.0 * b[i] + b[i + 1]) * 1.0;
    }
    a[n - 1] = 0.0;
    c[n - 1] = 3.0 * a[n - 2] * 1.0 * 1.0 + 2.0 * b[n - 2] * 1.0 + c[n - 2];
}

double spline_eval(const double* a, const double* b, const double* c, const double* y, size_t k, double x, bool cubic)
{
    const double h = x - double(k);
    return cubic ? ((a[k] * h + b[k]) * h + c[k]) * h + y[k] : c[k] * h + y[k];
}

// Hankel W-kernel cache of one plane: generate_convolution_kernel_wproj's Hankel
// branch (aw_projection.cpp:211-287), RadialInterpolate (aw_projection.h:127-213,
// non-diagonal) and generate_kernel_cache's Hankel branch (aw_projection.cpp:380-411).
// proj_slice: the 1-D projection of the 2-D image-domain kernel
// (generate_projected_image_domain_kernel, aw_projection.cpp:113-171) and its forward
// DFT instead of the radial profile and the DHT.
std::vector<std::vector<cx>> hankel_kernel_cache(double w_value, const std::vector<double>& aa, unsigned workarea_size,
    double cell_rad, unsigned oversampling, double scaling_factor, unsigned max_support, bool proj_slice, bool cubic,
    double trunc_perc, unsigned* trunc_support, std::vector<cx>* radius_out)
{
    const size_t arr = size_t(workarea_size) * oversampling;
    const size_t half_arr = arr / 2, half_k = workarea_size / 2;
    const double scaled_cell = cell_rad * scaling_factor;
    std::vector<cx> y;
    if (!proj_slice) {
        std::vector<cx> prof(half_arr, cx(0.0, 0.0));
        for (size_t i = 0; i < half_k; i++) {
            const double dx = double(i) * scaled_cell;
            const double r2 = dx * dx;
            cx t;
            if (r2 >= 1.0) {
                t = cx(1.0, 0.0);
            } else {
                const double nn = std::sqrt(1.0 - r2);
                const double ph = (-2.0 * M_PI) * w_value * (nn - 1.0);
                t = cx(std::cos(ph), std::sin(ph)) / nn;
            }
            prof[i] = t * aa[half_k + i];
        }
        static std::map<size_t, std::vector<double>> dht_cache;
        auto it = dht_cache.find(half_arr);
        if (it == dht_cache.end())
            it = dht_cache.emplace(half_arr, dht_matrix(half_arr)).first;
        const std::vector<double>& D = it->second;
        y.resize(half_arr);
        for (size_t i = 0; i < half_arr; i++) {
            double a1 = 0.0, a2 = 0.0;
            for (size_t j = 0; j < half_arr; j++) {
                a1 += D[j + i * half_arr] * prof[j].real();
                a2 += D[j + i * half_arr] * prof[j].imag();
            }
            y[i] = cx(a1, a2);
        }
    } else {
        std::vector<cx> proj(arr, cx(0.0, 0.0));
        for (size_t i = 1; i < half_k; i++)
            proj[i] = combine_2d_kernel_value(aa, i, 0, w_value, scaled_cell, unsigned(half_k));
        cx s0(0.0, 0.0);
        for (size_t i = 0; i < arr; i++)
            s0 += proj[i];
        proj[0] = s0 * cx(2.0, 0.0);
        proj[0] += aa[half_k] * aa[half_k];
        for (size_t j = 1; j < half_k; j++)
            for (size_t i = 1; i < half_k; i++)
                proj[j] += combine_2d_kernel_value(aa, i, j, w_value, scaled_cell, unsigned(half_k)) * cx(2.0, 0.0);
        for (size_t i = 1; i < half_k; i++)
            proj[arr - i] = proj[i];
        // forward DFT (fft_fftw_dft_c2c_1d, FFTW_FORWARD)
        y.resize(arr);
        for (size_t k = 0; k < arr; k++) {
            cx acc(0.0, 0.0);
            for (size_t j = 0; j < arr; j++) {
                const double ang = -2.0 * M_PI * double((j * k) % arr) / double(arr);
                acc += proj[j] * cx(std::cos(ang), std::sin(ang));
            }
            y[k] = acc;
        }
    }
    if (radius_out)
        *radius_out = y;
    const size_t n = y.size();
    // truncation (aw_projection.h:132-148)
    unsigned ts = max_support;
    if (trunc_perc > 0.0) {
        const double min_value = std::abs(y[0]) * (trunc_perc / 100.0);
        size_t idx = size_t(max_support) * oversampling;
        const double scale = 1.0 / double(oversampling);
        while (idx > oversampling) {
            if (std::abs(y[idx]) > min_value)
                break;
            idx -= oversampling;
        }
        ts = std::min(unsigned(std::ceil(double(idx) * scale)), max_support);
    }
    if (trunc_support)
        *trunc_support = ts;
    // RadialInterpolate: the half quadrant {(ii, j): ii <= j < ko} row by row
    const size_t ko = (size_t(ts) * 2 + 2) * oversampling / 2;
    std::vector<double> yr(n), yi(n), ar(n), br(n), cr(n), ai(n), bi(n), ci(n);
    for (size_t i = 0; i < n; i++) {
        yr[i] = y[i].real();
        yi[i] = y[i].imag();
    }
    spline_coefs(yr.data(), n, cubic, ar.data(), br.data(), cr.data());
    spline_coefs(yi.data(), n, cubic, ai.data(), bi.data(), ci.data());
    std::vector<cx> kh(ko + ko * (ko - 1) / 2);
    size_t q = 0;
    for (size_t ii = 0; ii < ko; ii++)
        for (size_t j = ii; j < ko; j++) {
            const double nx = (ii == 0) ? double(j) : std::sqrt(double(ii * ii + j * j));
            size_t k = size_t(nx);
            if (k >= n)
                k = n - 1;
            kh[q++] = cx(spline_eval(ar.data(), br.data(), cr.data(), yr.data(), k, nx, cubic),
                spline_eval(ai.data(), bi.data(), ci.data(), yi.data(), k, nx, cubic));
        }
    // generate_kernel_cache, Hankel branch
    const long long os = oversampling, op = os / 2, S = ts;
    const size_t cache_size = size_t(2 * op + 1), tcs = size_t(2 * S + 1);
    std::vector<std::vector<cx>> cache(cache_size * cache_size, std::vector<cx>(tcs * tcs));
    for (long long x = 0; x < (long long)cache_size; x++)
        for (long long yv = 0; yv < (long long)cache_size; yv++) {
            std::vector<cx>& tmp = cache[size_t(yv) * cache_size + size_t(x)];
            double ksum = 0.0;
            size_t ty = 0;
            for (long long cxv = -S * os + op - x; cxv < S * os + op - x + 1; cxv += os, ty++) {
                size_t tx = 0;
                for (long long cyv = -S * os + op - yv; cyv < S * os + op - yv + 1; cyv += os, tx++) {
                    const size_t dcx = size_t(cxv < 0 ? -cxv : cxv), dcy = size_t(cyv < 0 ? -cyv : cyv);
                    const size_t vi = dcx > dcy ? dcy * ko - ((dcy * (dcy + 1)) >> 1) + dcx
                                                : dcx * ko - ((dcx * (dcx + 1)) >> 1) + dcy;
                    tmp[tx + ty * tcs] = kh[vi];
                    ksum += kh[vi].real();
                }
            }
            for (auto& v : tmp)
                v = cx(v.real() / ksum, v.imag() / ksum);
        }
    return cache;
}

GridderOutput convolve_to_grid(const KernelFunc& kf, unsigned support, int image_size,
    const double* uv_in, const cx* vis_in, const double* weights_in, size_t n,
    bool kernel_exact, unsigned oversampling, bool shift_uv, bool halfplane, bool generate_beam,
    const WProjPars& wp, const double* w_in, double cell_size_arcsec, bool analytic_gcf, int nthreads)
{
    const int N = image_size;
    const int half = N / 2;
    int conv_support = int(support);
    const bool use_wproj = wp.num_wplanes > 0;
    if ((N % 2) != 0)
        throw std::invalid_argument("image_size must be even");
    if (use_wproj) {
        conv_support = int(wp.max_wpconv_support);
        kernel_exact = false;
    }
    if (conv_support <= 0)
        throw std::invalid_argument("support must be > 0");
    if (kernel_exact) {
        oversampling = 1;
    } else {
        oversampling = (oversampling >> 1) << 1;
        if (oversampling == 0)
            oversampling = 1;
    }

    // own copies (the reference takes these by value, gridder.h:245-253)
    std::vector<double> uv(uv_in, uv_in + 2 * n);
    std::vector<cx> vis(vis_in, vis_in + n);
    std::vector<double> wts(weights_in, weights_in + n);
    std::vector<double> wl(n, 0.0);
    if (use_wproj)
        wl.assign(w_in, w_in + n);
    std::vector<uint32_t> good(n, 1);

    if (use_wproj) {
        // sort by |w| ascending (gridder.h:324); stable for determinism
        std::vector<size_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return std::fabs(wl[a]) < std::fabs(wl[b]); });
        std::vector<double> uv2(2 * n), wl2(n), w2(n);
        std::vector<cx> vis2(n);
        for (size_t i = 0; i < n; i++) {
            uv2[i] = uv[idx[i]];
            uv2[n + i] = uv[n + idx[i]];
            wl2[i] = wl[idx[i]];
            vis2[i] = vis[idx[i]];
            w2[i] = wts[idx[i]];
        }
        uv.swap(uv2);
        wl.swap(wl2);
        vis.swap(vis2);
        wts.swap(w2);
    }
    if (halfplane)
        convert_to_halfplane_visibilities(uv, n, use_wproj ? &wl : nullptr, vis, conv_support, good);

    std::vector<int> kc(2 * n);
    std::vector<double> frac(2 * n);
    for (size_t i = 0; i < n; i++) {
        int val = int(std::rint(uv[i]));
        frac[i] = uv[i] - val;
        kc[i] = val + half;
        val = int(std::rint(uv[n + i]));
        frac[n + i] = uv[n + i] - val;
        kc[n + i] = val + half;
    }
    bounds_check_kernel_centre_locations(good, kc, n, N, conv_support);
    const int rows = halfplane ? half + 1 : N;
    if (shift_uv) {
        for (size_t i = 0; i < 2 * n; i++)
            kc[i] = (kc[i] < half) ? kc[i] + half : kc[i] - half;
    }
    double total = 0;
    size_t ngood = 0;
    for (size_t i = 0; i < n; i++)
        if (good[i] != 0) {
            total += wts[i];
            ngood++;
        }
    total *= 2.0;

    GridderOutput out;
    out.rows = rows;
    out.cols = N;
    out.sample_grid_total = total;
    out.num_good_vis = ngood;
    out.vis_grid.assign(size_t(rows) * N, cx(0.0, 0.0));
    if (generate_beam)
        out.sampling_grid.assign(size_t(rows) * N, cx(0.0, 0.0));
    cx* vg = out.vis_grid.data();
    cx* sg = generate_beam ? out.sampling_grid.data() : nullptr;

    if (!kernel_exact) {
        unsigned num_planes = 1;
        std::vector<double> w_avg;
        std::vector<size_t> first_idx(1, 0);
        std::vector<double> aa;
        unsigned workarea = unsigned(N);
        double scaling = 1.0;
        // cache[(cp_y * cs + cp_x)] -> (2s+1)^2 column-major, row = y tap
        std::vector<std::vector<cx>> cache;
        if (use_wproj) {
            num_planes = wp.num_wplanes;
            std::vector<double> absw(n);
            for (size_t i = 0; i < n; i++)
                absw[i] = std::fabs(wl[i]);
            average_w_planes(absw, good, num_planes, w_avg, first_idx, wp.wplanes_median);
            const int ks = 2 * conv_support + 1;
            const int us = wp.undersampling_opt > 0 ? int(std::pow(2, wp.undersampling_opt - 1)) : 0;
            if (us > 0) {
                workarea = 2;
                while (int(workarea) / us < ks)
                    workarea *= 2;
            } else {
                workarea = 2;
                while (int(workarea) < N)
                    workarea *= 2;
            }
            scaling = double(N) / double(workarea);
            aa = img_dom_kernel(kf, workarea, false, analytic_gcf);
        } else {
            cache = populate_kernel_cache(kf, support, oversampling);
        }
        const unsigned cache_size = (oversampling / 2) * 2 + 1;
        std::vector<int> ovs(2 * n);
        calculate_oversampled_kernel_indices(frac.data(), 2 * n, oversampling, ovs.data());
        for (auto& v : ovs)
            v += int(oversampling / 2);
        const double cell_rad = ((cell_size_arcsec / 3600.0) * (M_PI / 180.0));

        for (unsigned pi = 0; pi < num_planes; pi++) {
            const size_t vb = first_idx[pi];
            const size_t ve = (pi == num_planes - 1) ? n : first_idx[pi + 1];
            int cs_plane = conv_support;
            if (use_wproj) {
                unsigned ts = 0;
                if (wp.hankel)
                    cache = hankel_kernel_cache(w_avg[pi], aa, workarea, cell_rad, oversampling, scaling, unsigned(conv_support),
                        wp.hankel_proj_slice, wp.interp_cubic, wp.kernel_trunc_perc, &ts, nullptr);
                else
                    cache = wproj_kernel_cache(w_avg[pi], aa, workarea, cell_rad, oversampling, scaling, unsigned(conv_support),
                        nthreads, wp.kernel_trunc_perc, &ts);
                // the plane grids with its truncated support (gridder.h:756-758)
                cs_plane = int(ts);
            }
            const int ks = 2 * cs_plane + 1;
            // column-owner parallelism: per pixel, contributions add in visibility order,
            // exactly as each TBB column-residue task does (gridder.h:767-934)
            const int T = std::max(1, std::min(ks, nthreads));
            auto worker = [&](int t) {
                for (size_t vi = vb; vi < ve; vi++) {
                    const uint32_t g = good[vi];
                    if (g == 0)
                        continue;
                    int gc_x = kc[vi], gc_y = kc[n + vi];
                    int cp_x = ovs[vi], cp_y = ovs[n + vi];
                    cx vv = vis[vi];
                    const double wt = wts[vi];
                    double wlv = wl[vi];
                    for (uint32_t gv = 0; gv < g; gv++) {
                        if (gv) {
                            gc_x = -kc[vi] + N;
                            gc_y = -kc[n + vi] + N;
                            if (oversampling > 1) {
                                cp_x = -ovs[vi] + int(oversampling);
                                cp_y = -ovs[n + vi] + int(oversampling);
                            }
                            vv = std::conj(vv);
                            wlv *= -1;
                        }
                        const std::vector<cx>& K = cache[size_t(cp_y) * cache_size + cp_x];
                        const bool conjk = use_wproj && (wlv < 0.0);
                        for (int d = -cs_plane; d <= cs_plane; d++) {
                            int col = gc_x + d;
                            col = ((col % N) + N) % N;
                            if ((col % T) != t)
                                continue;
                            const cx* kcol = &K[size_t(cs_plane + d) * ks];
                            int r = gc_y - cs_plane;
                            for (int i = 0; i < ks; i++, r++) {
                                if (r < 0)
                                    continue;
                                int rr = r;
                                if (rr >= N)
                                    rr -= N;
                                if (rr < rows) {
                                    const cx kv0 = conjk ? std::conj(kcol[i]) : kcol[i];
                                    const cx kval(kv0.real() * wt, kv0.imag() * wt);
                                    vg[rr + size_t(col) * rows] += vv * kval;
                                    if (sg)
                                        sg[rr + size_t(col) * rows] += kval;
                                }
                            }
                        }
                    }
                }
            };
            if (T == 1) {
                worker(0);
            } else {
                std::vector<std::thread> th;
                for (int t = 0; t < T; t++)
                    th.emplace_back(worker, t);
                for (auto& x : th)
                    x.join();
            }
        }
    } else {
        // exact gridder (gridder.h:943-994)
        const int ks = 2 * conv_support + 1;
        for (size_t vi = 0; vi < n; vi++) {
            if (good[vi] == 0)
                continue;
            int gc_x = kc[vi], gc_y = kc[n + vi];
            double fx = frac[vi], fy = frac[n + vi];
            cx vv = vis[vi];
            const double wt = wts[vi];
            for (uint32_t gv = 0; gv < good[vi]; gv++) {
                if (gv == 1) {
                    gc_x = -kc[vi] + N;
                    gc_y = -kc[n + vi] + N;
                    fx *= -1;
                    fy *= -1;
                    vv = std::conj(vv);
                }
                const std::vector<double> K = make_kernel_array(kf, conv_support, fx, fy, 1, false, true);
                for (int j = 0; j < ks; j++) {
                    int col = gc_x - conv_support + j;
                    if (col < 0)
                        col += N;
                    if (col >= N)
                        col -= N;
                    for (int i = 0; i < ks; i++) {
                        int r = gc_y - conv_support + i;
                        if (r < 0)
                            r += N;
                        if (r >= N)
                            r -= N;
                        if (r < rows) {
                            const double kv = wt * K[i + size_t(j) * ks];
                            const cx kval(kv, 0.0);
                            vg[r + size_t(col) * rows] += vv * kval;
                            if (sg)
                                sg[r + size_t(col) * rows] += kval;
                        }
                    }
                }
            }
        }
    }
    return out;
}

} // namespace oracle
