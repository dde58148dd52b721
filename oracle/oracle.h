/*
 * ORACLE - TEST INFRASTRUCTURE ONLY.
 *
 * A dependency-free FP64 CPU restatement of the reference Slow Transients
 * Pipeline imaging + source-finding path (SKA-ScienceDataProcessor/FastImaging,
 * USE_FLOAT=OFF, FFTSHIFT=OFF build).  It is the checker the HIP path is
 * compared against and the CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference repository root).  Deliberate, documented deviations:
 *  - Sums whose order the reference leaves to TBB (mean/sigma, sigma-clip sums)
 *    use one fixed "canonical" order (canonical_pair_sum below) that the HIP
 *    kernels reproduce exactly, so statistics and thresholds are bit-identical.
 *  - 8-connected labelling uses the partition-independent definition (every
 *    visual 8-neighbour link, no wrap); see ccl_label() for the one link the
 *    reference's slab scan drops at interior slab columns.
 *  - Extremum ties resolve to the smallest memory index.
 *  - The 2-D FFTs are this file's own radix-2 transforms with FFTW's
 *    semantics (FFTW is absent from this environment).
 */
#ifndef STP_ORACLE_H
#define STP_ORACLE_H

#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../include/stp_capi.h"

namespace oracle {

using cx = std::complex<double>;

/* ---------------- kernel functors: src/stp/convolution/conv_func.{h,cpp} ---------------- */
struct KernelFunc {
    int kind;        // stp_kernel_function
    double param;    // TopHat/Triangle: half_base_width; Sinc/Gaussian/GaussianSinc/PSWF: trunc
    double width_g;  // Gaussian width normalisation
    double width_s;  // Sinc width normalisation
};
KernelFunc make_kernel_func(int kind, double param);
// From a kernel spec (functor arguments as given, conv_func.h:23-262).
KernelFunc make_kernel_func(const stp_kernel_spec& spec);
// Applies the functor element-wise (operator() of the functor classes).
void eval_kernel(const KernelFunc& k, const double* x, size_t n, double* out);
// gcf() of the functor (analytic grid-correction); false if the reference asserts.
bool eval_gcf(const KernelFunc& k, const double* x, size_t n, double* out);

// make_1D_kernel (conv_func.h:319-337)
std::vector<double> make_1d_kernel(const KernelFunc& k, int support, double offset, int oversampling, bool pad, bool normalize);
// make_kernel_array (conv_func.h:279-303); result (2s+1)x(2s+1) column-major, rows = y.
std::vector<double> make_kernel_array(const KernelFunc& k, int support, double off_x, double off_y, int oversampling, bool pad, bool normalize);
// ImgDomKernel (conv_func.h:351-412)
std::vector<double> img_dom_kernel(const KernelFunc& k, size_t kernel_size, bool normalize, bool analytic_gcf);

/* ---------------- FFT: src/stp/common/fft.cpp semantics ---------------- */
// In-place 1-D complex FFT, sign -1 (forward) or +1 (inverse), unnormalised, n power of 2.
void fft1d(cx* data, size_t n, int sign);
// fft_fftw_c2r (fft.cpp:45-135): in (n/2+1) x cols complex col-major -> out n x cols real.
void fft_c2r_2d(const cx* in, int in_rows, int cols, double* out, int nthreads);
// fft_fftw_c2c (fft.cpp:227-323): n_rows x n_cols complex col-major, sign -1 forward.
void fft_c2c_2d(const cx* in, int rows, int cols, bool forward, cx* out, int nthreads);
// fft_fftw_dft_r2r_1d with FFTW_R2HC (fft.cpp:325-412).
void fft_r2hc_1d(const double* in, size_t n, double* out);

/* ---------------- gridder: src/stp/gridder/gridder.{h,cpp} ---------------- */
struct GridderOutput {
    std::vector<cx> vis_grid;      // rows x cols column-major
    std::vector<cx> sampling_grid; // rows x cols (empty if no beam)
    double sample_grid_total = 0.0;
    int rows = 0, cols = 0;
    size_t num_good_vis = 0;
};

std::vector<std::vector<cx>> populate_kernel_cache(const KernelFunc& kf, unsigned support, unsigned oversampling);
void convert_to_halfplane_visibilities(std::vector<double>& uv, size_t n, std::vector<double>* w, std::vector<cx>& vis, int support, std::vector<uint32_t>& good);
void bounds_check_kernel_centre_locations(std::vector<uint32_t>& good, const std::vector<int>& kc, size_t n, int image_size, int support);
void calculate_oversampled_kernel_indices(const double* frac, size_t n, unsigned oversampling, int* out);
void average_w_planes(const std::vector<double>& abs_w, const std::vector<uint32_t>& good, unsigned num_wplanes,
    std::vector<double>& w_avg, std::vector<size_t>& first_idx, bool median);

struct WProjPars {
    unsigned num_wplanes = 0, max_wpconv_support = 0, undersampling_opt = 1;
    double kernel_trunc_perc = 0.0;
    bool hankel = false, wplanes_median = false;
    bool hankel_proj_slice = false, interp_cubic = false; // Hankel kernels: projection slice; cubic spline (interp_type 2)
};

GridderOutput convolve_to_grid(const KernelFunc& kf, unsigned support, int image_size,
    const double* uv, const cx* vis, const double* weights, size_t n,
    bool kernel_exact, unsigned oversampling, bool shift_uv, bool halfplane, bool generate_beam,
    const WProjPars& wp, const double* w_lambda, double cell_size_arcsec, bool analytic_gcf, int nthreads);

// W-projection kernel cache for one plane (aw_projection.cpp:50-111,173-210,366-452).
// Returns cache[(y*cache_size + x)] of (2s+1)^2 complex each, column-major (row = y tap).
// With trunc_perc > 0 the support is first truncated (aw_projection.cpp:197-209); the
// support used is returned in *trunc_support (may be NULL).
std::vector<std::vector<cx>> wproj_kernel_cache(double w_value, const std::vector<double>& aa_kernel_img,
    unsigned workarea_size, double cell_rad, unsigned oversampling, double scaling_factor, unsigned support, int nthreads,
    double trunc_perc = 0.0, unsigned* trunc_support = nullptr);

// Hankel W-kernels (aw_projection.cpp:211-287,380-411,477-596; aw_projection.h:127-213).
double besselj1(double x);
std::vector<double> dht_matrix(size_t n);
void spline_coefs(const double* y, size_t n, bool cubic, double* a, double* b, double* c);
double spline_eval(const double* a, const double* b, const double* c, const double* y, size_t k, double x, bool cubic);
// Cache of one plane, laid out as wproj_kernel_cache's; radius_out (may be NULL)
// receives the transformed radial profile (DHT) or the DFT of the projection.
std::vector<std::vector<cx>> hankel_kernel_cache(double w_value, const std::vector<double>& aa_kernel_img,
    unsigned workarea_size, double cell_rad, unsigned oversampling, double scaling_factor, unsigned max_support,
    bool proj_slice, bool cubic, double trunc_perc, unsigned* trunc_support, std::vector<cx>* radius_out);

/* ---------------- imager: src/stp/imager/imager.h ---------------- */
struct ImagerResult {
    std::vector<double> image, beam; // image_size x image_size col-major (empty if total <= 0)
    double sample_grid_total = 0.0;
    size_t num_good_vis = 0;
};
ImagerResult image_visibilities(const stp_imager_pars& p, const stp_wproj_pars* wp, const stp_kernel_spec* kernel, size_t n,
    const cx* vis, const double* weights, const double* uvw, int nthreads);

/* ---------------- statistics: src/stp/common/matrix_math.cpp ---------------- */
struct Stats {
    double mean = 0, sigma = 0, median = 0;
    bool mean_valid = false, sigma_valid = false, median_valid = false;
};
// Fixed-order sums of x and x*x (see definition in stats.cpp).
void canonical_pair_sum(const double* x, size_t n, double* s1, double* s2);
Stats mat_mean_and_stddev(const double* x, size_t n);
double mat_median_exact(const double* x, size_t n);
Stats mat_binmedian(const double* x, size_t n);
Stats mat_median_binapprox(const double* x, size_t n);
double estimate_rms(const double* x, size_t n, double num_sigma, unsigned iters, Stats st);

/* ---------------- source find: src/stp/sourcefind + src/stp/common/ccl.h ---------------- */
// Provisional connected-component labelling.  Returns label map (int32, rows x cols)
// with +id / -id final ids (id = rank of the component's first pixel in the scan order).
void ccl_label(const double* img, int rows, int cols, double thr_pos, double thr_neg,
    bool find_negative, bool conn4, int32_t* labels, uint32_t* n_pos, uint32_t* n_neg);
// Restatement of the reference slab-partitioned SAUF scan (ccl.h:107-484, :505-773)
// with slabs of `slab_cols` visual columns; used to compare partition effects.
void ccl_label_sauf(const double* img, int rows, int cols, double thr_pos, double thr_neg,
    bool find_negative, bool conn4, int slab_cols, int32_t* labels, uint32_t* n_pos, uint32_t* n_neg);

struct SourceFindResult {
    std::vector<int32_t> label_map;
    std::vector<stp_island> islands;
    stp_sourcefind_stats stats{};
};
SourceFindResult source_find(const double* img, int rows, int cols, const stp_sourcefind_pars& p);
void estimate_moments_fit(stp_island& isl, double x_bar, double y_bar, double xx_bar, double yy_bar, double xy_bar,
    double rms_est, double analysis_n_sigma);

/* ---------------- model visibilities: src/stp/visibility/visibility.cpp ---------------- */
void model_visibilities(const double* skymodel, size_t n_src, const double* uvw, size_t n_vis, cx* out, int nthreads);

/* ---------------- island fit: src/stp/sourcefind/sourcefind.cpp:719-866, fitting.cpp ---------------- */
constexpr double LM_MU0 = STP_FIT_MU0, LM_MU_MAX = STP_FIT_MU_MAX, LM_DIAG_MIN = STP_FIT_DIAG_MIN,
                 LM_DIAG_MAX = STP_FIT_DIAG_MAX, LM_FTOL = STP_FIT_FTOL, LM_XTOL = STP_FIT_XTOL;
constexpr int LM_MAX_ITERS = STP_FIT_MAX_ITERS;
struct FitResult {
    stp_gaussian2d fit;   // constrained
    double half_cost = 0.0;
    int iterations = 0;
    int status = 0;
};
// Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57)
void fit_constrain(stp_gaussian2d& g);
// Residual model - data at (x, y) and its Jacobian row (fitting.cpp:59-120); p = (A, x0, y0, sx, sy, theta).
void gaussian_jacobian(const double* p, double x, double y, double v, double* r, double* J);
// Least-squares fit of m samples from `start` (unconstrained parameters).
FitResult lm_fit(const double* xs, const double* ys, const double* vs, size_t m, const double* start);
// Fits of every island of a source find (label map and image: rows x cols, column-major).
std::vector<stp_fit_result> fit_islands(const double* data, int rows, int cols, const int32_t* labels,
    const stp_island* islands, size_t n);

int default_threads();

} // namespace oracle

#endif
