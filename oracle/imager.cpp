/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Imager: src/stp/imager/imager.h:31-214 (non-FFTSHIFT normalisation branch
 * :96-213) and :233-376 (image_visibilities), dispatch imager.cpp:13-69.
 */
#include <cmath>
#include <stdexcept>

#include "oracle.h"

namespace oracle {

// normalise_image_beam_result_1D, non-FFTSHIFT branch (imager.h:96-213):
// crops the padded quadrants into the image_size output, keeping the
// un-shifted layout, scales by the normalisation factor and optionally divides
// by the grid-correction function.
static void normalise(const double* img, size_t padded, size_t image_size, double factor,
    const std::vector<double>* gcf, double* out)
{
    const size_t hp = padded / 2, hi = image_size / 2;
    auto put = [&](size_t i, size_t j, size_t oi, size_t oj, size_t ii, size_t jj) {
        const double v = img[oi + oj * padded];
        out[i + j * image_size] = gcf ? v * factor / ((*gcf)[ii] * (*gcf)[jj]) : v * factor;
    };
    for (size_t j = 0; j < hi; j++) {
        const size_t jj = j + hp;
        size_t ii = hp;
        for (size_t i = 0; i < hi; i++, ii++)
            put(i, j, i, j, ii, jj);
        ii = hp - hi;
        size_t oi = padded - hi;
        for (size_t i = hi; i < image_size; i++, ii++, oi++)
            put(i, j, oi, j, ii, jj);
    }
    for (size_t j = hi; j < image_size; j++) {
        const size_t oj = padded + j - image_size;
        const size_t jj = hp + j - image_size;
        size_t ii = hp;
        for (size_t i = 0; i < hi; i++, ii++)
            put(i, j, i, oj, ii, jj);
        ii = hp - hi;
        size_t oi = padded - hi;
        for (size_t i = hi; i < image_size; i++, ii++, oi++)
            put(i, j, oi, oj, ii, jj);
    }
}

ImagerResult image_visibilities(const stp_imager_pars& p, const stp_wproj_pars* wpp, const stp_kernel_spec* kernel, size_t n,
    const cx* vis, const double* weights, const double* uvw, int nthreads)
{
    const int padded = int(p.padded_image_size);
    if (p.padding_factor < 1.0 || padded < int(p.image_size) || padded <= 0 || (padded & (padded - 1)))
        throw std::invalid_argument("padded image size must be a power of two >= image_size");
    if (p.kernel_support == 0 || !(p.cell_size > 0.0))
        throw std::invalid_argument("kernel_support and cell_size must be > 0");
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
        if (p.kernel_exact)
            throw std::invalid_argument("W-projection requires kernel_exact == false (imager.h:287)");
    }
    // kernel_creator: given, or built by the ImageVisibilities dispatch with
    // trunc / half_base_width = kernel_support (imager.cpp:26-60)
    const KernelFunc kf = kernel ? make_kernel_func(*kernel) : make_kernel_func(p.kernel_function, double(p.kernel_support));

    // uv -> pixels (imager.h:292-300)
    const double inv_px = ((p.cell_size / 3600.0) * (M_PI / 180.0)) * double(padded);
    std::vector<double> uv(2 * n), w(n);
    for (size_t i = 0; i < n; i++) {
        uv[i] = uvw[i] * inv_px;
        uv[n + i] = uvw[n + i] * inv_px;
        w[i] = uvw[2 * n + i];
    }
    GridderOutput g = convolve_to_grid(kf, p.kernel_support, padded, uv.data(), vis, weights, n,
        p.kernel_exact != 0, p.oversampling, true, true, p.generate_beam != 0, wp, w.data(), p.cell_size,
        p.analytic_gcf != 0, nthreads);

    ImagerResult res;
    res.sample_grid_total = g.sample_grid_total;
    res.num_good_vis = g.num_good_vis;
    std::vector<double> img(size_t(padded) * padded), beam;
    fft_c2r_2d(g.vis_grid.data(), g.rows, g.cols, img.data(), nthreads);
    if (p.generate_beam) {
        beam.resize(size_t(padded) * padded);
        fft_c2r_2d(g.sampling_grid.data(), g.rows, g.cols, beam.data(), nthreads);
    }
    if (g.sample_grid_total > 0.0) {
        const double factor = 1.0 / g.sample_grid_total;
        std::vector<double> gcf;
        if (p.gridding_correction)
            gcf = img_dom_kernel(kf, size_t(padded), false, p.analytic_gcf != 0);
        const size_t isz = p.image_size;
        res.image.resize(isz * isz);
        normalise(img.data(), size_t(padded), isz, factor, p.gridding_correction ? &gcf : nullptr, res.image.data());
        if (p.generate_beam) {
            res.beam.resize(isz * isz);
            normalise(beam.data(), size_t(padded), isz, factor, p.gridding_correction ? &gcf : nullptr, res.beam.data());
        }
    }
    return res;
}

} // namespace oracle
