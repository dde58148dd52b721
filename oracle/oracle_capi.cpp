/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * extern "C" surface of the oracle for the Python test harness (ctypes).
 */
#include <cstring>
#include <exception>
#include <string>

#include "oracle.h"

using namespace oracle;

static thread_local std::string g_err;

template <typename F>
static int guard(F f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

int oracle_model_visibilities(const double* skymodel, size_t n_src, const double* uvw, size_t n_vis, double* out, int nthreads)
{
    return guard([&] { model_visibilities(skymodel, n_src, uvw, n_vis, reinterpret_cast<cx*>(out), nthreads); });
}
int oracle_default_threads() { return default_threads(); }

int oracle_kernel_eval(const stp_kernel_spec* k, const double* x, size_t n, double* out)
{
    return guard([&] { eval_kernel(make_kernel_func(*k), x, n, out); });
}

int oracle_make_1d_kernel(const stp_kernel_spec* k, int support, double offset, int os, int pad, int normalize, double* out)
{
    return guard([&] {
        auto v = make_1d_kernel(make_kernel_func(*k), support, offset, os, pad != 0, normalize != 0);
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

int oracle_make_kernel_array(const stp_kernel_spec* k, int support, double off_x, double off_y, int os, int pad, int normalize, double* out)
{
    return guard([&] {
        auto v = make_kernel_array(make_kernel_func(*k), support, off_x, off_y, os, pad != 0, normalize != 0);
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

int oracle_img_dom_kernel(const stp_kernel_spec* k, size_t size, int normalize, int analytic, double* out)
{
    return guard([&] {
        auto v = img_dom_kernel(make_kernel_func(*k), size, normalize != 0, analytic != 0);
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

// out: cs x cs blocks of (2s+1)^2 complex, block (a, b) at offset (a*cs + b)
int oracle_populate_kernel_cache(const stp_kernel_spec* k, unsigned support, unsigned os, double* out)
{
    return guard([&] {
        auto c = populate_kernel_cache(make_kernel_func(*k), support, os);
        size_t off = 0;
        for (auto& k : c) {
            std::memcpy(out + off, k.data(), k.size() * sizeof(cx));
            off += 2 * k.size();
        }
    });
}

int oracle_oversampled_indices(const double* frac, size_t n, unsigned os, int* out)
{
    return guard([&] { calculate_oversampled_kernel_indices(frac, n, os, out); });
}

int oracle_convolve_to_grid(const stp_kernel_spec* k, unsigned support, int image_size, size_t n,
    const double* uv, const double* vis, const double* w, int exact, unsigned os, int shift_uv, int halfplane,
    int beam, const stp_wproj_pars* wpp, const double* w_lambda, double cell_size, int analytic_gcf, int nthreads,
    double* vis_grid, double* sampling_grid, double* total, uint64_t* ngood)
{
    return guard([&] {
        WProjPars wp;
        if (wpp && wpp->num_wplanes > 0) {
            wp.num_wplanes = wpp->num_wplanes;
            wp.max_wpconv_support = wpp->max_wpconv_support;
            wp.undersampling_opt = wpp->undersampling_opt;
            wp.kernel_trunc_perc = wpp->kernel_trunc_perc;
            wp.hankel = wpp->hankel_opt != 0;
            wp.hankel_proj_slice = wpp->hankel_proj_slice != 0;
            wp.interp_cubic = wpp->interp_type == 2;
            wp.wplanes_median = wpp->wplanes_median != 0;
        }
        GridderOutput g = convolve_to_grid(make_kernel_func(*k), support, image_size, uv,
            reinterpret_cast<const cx*>(vis), w, n, exact != 0, os, shift_uv != 0, halfplane != 0, beam != 0, wp, w_lambda,
            cell_size, analytic_gcf != 0, nthreads);
        std::memcpy(vis_grid, g.vis_grid.data(), g.vis_grid.size() * sizeof(cx));
        if (beam && sampling_grid)
            std::memcpy(sampling_grid, g.sampling_grid.data(), g.sampling_grid.size() * sizeof(cx));
        *total = g.sample_grid_total;
        if (ngood)
            *ngood = g.num_good_vis;
    });
}

int oracle_fft_c2r(const double* in, int in_rows, int cols, double* out, int nthreads)
{
    return guard([&] { fft_c2r_2d(reinterpret_cast<const cx*>(in), in_rows, cols, out, nthreads); });
}

int oracle_fft_c2c(const double* in, int rows, int cols, int forward, double* out, int nthreads)
{
    return guard([&] { fft_c2c_2d(reinterpret_cast<const cx*>(in), rows, cols, forward != 0, reinterpret_cast<cx*>(out), nthreads); });
}

int oracle_wproj_kernel_cache(double w_value, const double* aa, unsigned workarea, double cell_rad, unsigned os,
    double scaling, unsigned support, int nthreads, double trunc_perc, unsigned* trunc_support, double* out)
{
    return guard([&] {
        std::vector<double> a(aa, aa + workarea);
        auto c = wproj_kernel_cache(w_value, a, workarea, cell_rad, os, scaling, support, nthreads, trunc_perc, trunc_support);
        size_t off = 0;
        for (auto& k : c) {
            std::memcpy(out + off, k.data(), k.size() * sizeof(cx));
            off += 2 * k.size();
        }
    });
}

int oracle_hankel_kernel_cache(double w_value, const double* aa, unsigned workarea, double cell_rad, unsigned os,
    double scaling, unsigned support, int proj_slice, int cubic, double trunc_perc, unsigned* trunc_support, double* out,
    double* radius, size_t radius_cap, size_t* radius_n)
{
    return guard([&] {
        std::vector<double> a(aa, aa + workarea);
        std::vector<cx> rad;
        auto c = hankel_kernel_cache(w_value, a, workarea, cell_rad, os, scaling, support, proj_slice != 0, cubic != 0,
            trunc_perc, trunc_support, &rad);
        size_t off = 0;
        for (auto& k : c) {
            std::memcpy(out + off, k.data(), k.size() * sizeof(cx));
            off += 2 * k.size();
        }
        if (radius_n)
            *radius_n = rad.size();
        if (radius)
            std::memcpy(radius, rad.data(), std::min(radius_cap, rad.size()) * sizeof(cx));
    });
}

int oracle_besselj1(const double* x, size_t n, double* out)
{
    return guard([&] {
        for (size_t i = 0; i < n; i++)
            out[i] = besselj1(x[i]);
    });
}

int oracle_dht_matrix(size_t n, double* out)
{
    return guard([&] {
        const std::vector<double> d = dht_matrix(n);
        std::memcpy(out, d.data(), d.size() * sizeof(double));
    });
}

int oracle_spline(const double* y, size_t n, int cubic, const double* x, size_t m, double* out)
{
    return guard([&] {
        if (n < 3)
            throw std::invalid_argument("spline needs at least 3 points");
        std::vector<double> a(n), b(n), c(n);
        spline_coefs(y, n, cubic != 0, a.data(), b.data(), c.data());
        for (size_t i = 0; i < m; i++) {
            size_t k = x[i] < 0.0 ? 0 : size_t(x[i]);
            if (k >= n)
                k = n - 1;
            out[i] = spline_eval(a.data(), b.data(), c.data(), y, k, x[i], cubic != 0);
        }
    });
}

int oracle_average_w_planes(const double* abs_w, const uint32_t* good, size_t n, unsigned planes, int median, double* w_avg,
    uint64_t* first_idx)
{
    return guard([&] {
        std::vector<double> w(abs_w, abs_w + n);
        std::vector<uint32_t> g(good, good + n);
        std::vector<double> avg;
        std::vector<size_t> first;
        average_w_planes(w, g, planes, avg, first, median != 0);
        for (unsigned p = 0; p < planes; p++) {
            w_avg[p] = avg[p];
            first_idx[p] = first[p];
        }
    });
}

// Returns 0 with *empty = 1 when sample_grid_total <= 0 (outputs untouched).
int oracle_image_visibilities(const stp_imager_pars* p, const stp_wproj_pars* wp, const stp_kernel_spec* k, size_t n, const double* vis,
    const double* weights, const double* uvw, int nthreads, double* image, double* beam, double* total, uint64_t* ngood, int* empty)
{
    return guard([&] {
        ImagerResult r = image_visibilities(*p, wp, k, n, reinterpret_cast<const cx*>(vis), weights, uvw, nthreads);
        *total = r.sample_grid_total;
        if (ngood)
            *ngood = r.num_good_vis;
        *empty = r.image.empty() ? 1 : 0;
        if (!r.image.empty())
            std::memcpy(image, r.image.data(), r.image.size() * sizeof(double));
        if (beam && !r.beam.empty())
            std::memcpy(beam, r.beam.data(), r.beam.size() * sizeof(double));
    });
}

int oracle_canonical_pair_sum(const double* x, size_t n, double* out2)
{
    return guard([&] { canonical_pair_sum(x, n, &out2[0], &out2[1]); });
}

int oracle_mean_stddev(const double* x, size_t n, double* out2)
{
    return guard([&] {
        Stats s = mat_mean_and_stddev(x, n);
        out2[0] = s.mean;
        out2[1] = s.sigma;
    });
}

int oracle_median_exact(const double* x, size_t n, double* out)
{
    return guard([&] { *out = mat_median_exact(x, n); });
}

int oracle_binmedian(const double* x, size_t n, double* out3)
{
    return guard([&] {
        Stats s = mat_binmedian(x, n);
        out3[0] = s.mean;
        out3[1] = s.sigma;
        out3[2] = s.median;
    });
}

int oracle_binapprox(const double* x, size_t n, double* out3)
{
    return guard([&] {
        Stats s = mat_median_binapprox(x, n);
        out3[0] = s.mean;
        out3[1] = s.sigma;
        out3[2] = s.median;
    });
}

// stats3 = (mean, sigma, median); valid bits as flags (1 mean, 2 sigma, 4 median)
int oracle_estimate_rms(const double* x, size_t n, double num_sigma, unsigned iters, const double* stats3, int valid, double* out)
{
    return guard([&] {
        Stats s;
        s.mean = stats3[0];
        s.sigma = stats3[1];
        s.median = stats3[2];
        s.mean_valid = valid & 1;
        s.sigma_valid = valid & 2;
        s.median_valid = valid & 4;
        *out = estimate_rms(x, n, num_sigma, iters, s);
    });
}

int oracle_ccl(const double* img, int rows, int cols, double tp, double tn, int find_negative, int conn4,
    int32_t* labels, uint32_t* n_pos, uint32_t* n_neg)
{
    return guard([&] { ccl_label(img, rows, cols, tp, tn, find_negative != 0, conn4 != 0, labels, n_pos, n_neg); });
}

int oracle_ccl_sauf(const double* img, int rows, int cols, double tp, double tn, int find_negative, int conn4,
    int slab_cols, int32_t* labels, uint32_t* n_pos, uint32_t* n_neg)
{
    return guard([&] { ccl_label_sauf(img, rows, cols, tp, tn, find_negative != 0, conn4 != 0, slab_cols, labels, n_pos, n_neg); });
}

int oracle_source_find(const double* img, int rows, int cols, const stp_sourcefind_pars* p, int32_t* label_map,
    stp_island* islands, size_t cap, size_t* n_islands, stp_sourcefind_stats* stats)
{
    return guard([&] {
        SourceFindResult r = source_find(img, rows, cols, *p);
        if (label_map)
            std::memcpy(label_map, r.label_map.data(), r.label_map.size() * sizeof(int32_t));
        *n_islands = r.islands.size();
        for (size_t i = 0; i < r.islands.size() && i < cap; i++)
            islands[i] = r.islands[i];
        if (stats)
            *stats = r.stats;
    });
}

int oracle_gaussian_jacobian(const double* p, double x, double y, double v, double* r, double* J)
{
    return guard([&] { gaussian_jacobian(p, x, y, v, r, J); });
}

int oracle_lm_fit(const double* xs, const double* ys, const double* vs, size_t m, const double* start, double* fit6,
    double* half_cost, int* iterations, int* status)
{
    return guard([&] {
        const FitResult f = lm_fit(xs, ys, vs, m, start);
        const double v[6] = { f.fit.amplitude, f.fit.x_centre, f.fit.y_centre, f.fit.semimajor, f.fit.semiminor, f.fit.theta };
        std::memcpy(fit6, v, sizeof(v));
        *half_cost = f.half_cost;
        *iterations = f.iterations;
        *status = f.status;
    });
}

int oracle_fit_islands(const double* img, int rows, int cols, const int32_t* labels, const stp_island* islands, size_t n,
    stp_fit_result* out)
{
    return guard([&] {
        const std::vector<stp_fit_result> r = fit_islands(img, rows, cols, labels, islands, n);
        for (size_t k = 0; k < n; k++)
            out[k] = r[k];
    });
}

} // extern "C"
