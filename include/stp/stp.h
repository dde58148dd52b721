/*
 * stp/stp.h - C++ drop-in for the reference's stp:: imaging and source-finding
 * API (SKA-ScienceDataProcessor/FastImaging src/stp/stp.h), executed on an
 * MI355X through the C-ABI of include/stp_capi.h.
 *
 * Header-only.  Link with fastimaging_amd/lib/libstp_hip.so.  Armadillo is not
 * required: stp::Mat<T> is a column-major matrix with the members the
 * reference's callers use (n_rows, n_cols, n_elem, memptr(), colptr(), at(),
 * operator()); when <armadillo> is available, stp::Mat converts to and from
 * arma::Mat.
 *
 * Every call runs on a per-thread default context (device 0, or the device
 * set with stp::set_device()).  Errors are reported as std::runtime_error, as
 * the reference does for its runtime failures.
 */
#ifndef STP_STP_H
#define STP_STP_H

#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../stp_capi.h"

#if __has_include(<armadillo>)
#include <armadillo>
#define STP_HAVE_ARMADILLO 1
#endif

namespace stp {

using real_t = double;                  // the FP64 (USE_FLOAT=OFF) build, src/stp/types.h:16-18
using cx_real_t = std::complex<double>;
const real_t fptolerance = 1.0e-10;     // src/stp/types.h:26

/* ---------------------------------------------------------------- matrices */
template <typename T>
class Mat {
public:
    Mat() = default;
    Mat(size_t rows, size_t cols)
        : n_rows(rows)
        , n_cols(cols)
        , n_elem(rows * cols)
        , data_(rows * cols, T())
    {
    }
    void set_size(size_t rows, size_t cols)
    {
        n_rows = rows;
        n_cols = cols;
        n_elem = rows * cols;
        data_.assign(n_elem, T());
    }
    void reset() { set_size(0, 0); }
    bool is_empty() const { return n_elem == 0; }
    T* memptr() { return data_.data(); }
    const T* memptr() const { return data_.data(); }
    T* colptr(size_t c) { return data_.data() + c * n_rows; }
    const T* colptr(size_t c) const { return data_.data() + c * n_rows; }
    T& at(size_t r, size_t c) { return data_[r + c * n_rows]; }
    const T& at(size_t r, size_t c) const { return data_[r + c * n_rows]; }
    T& operator()(size_t r, size_t c) { return at(r, c); }
    const T& operator()(size_t r, size_t c) const { return at(r, c); }
    T& operator[](size_t i) { return data_[i]; }
    const T& operator[](size_t i) const { return data_[i]; }
#ifdef STP_HAVE_ARMADILLO
    Mat(const arma::Mat<T>& m)
        : Mat(m.n_rows, m.n_cols)
    {
        std::copy(m.memptr(), m.memptr() + n_elem, data_.begin());
    }
    operator arma::Mat<T>() const { return arma::Mat<T>(data_.data(), n_rows, n_cols); }
#endif
    size_t n_rows = 0, n_cols = 0, n_elem = 0;

private:
    std::vector<T> data_;
};
using mat = Mat<double>;
using cx_mat = Mat<std::complex<double>>;

/* ---------------------------------------------------- enums (types.h:32-88) */
enum struct KernelFunction { TopHat, Triangle, Sinc, Gaussian, GaussianSinc, PSWF };
enum struct FFTRoutine { FFTW_ESTIMATE_FFT, FFTW_MEASURE_FFT, FFTW_PATIENT_FFT, FFTW_WISDOM_FFT, FFTW_WISDOM_INPLACE_FFT };
enum struct InterpType { LINEAR, CUBIC, COSINE };
enum struct MedianMethod { ZEROMEDIAN, BINMEDIAN, BINAPPROX, NTHELEMENT };
enum struct CeresDiffMethod { AutoDiff, AutoDiff_SingleResBlk, AnalyticDiff, AnalyticDiff_SingleResBlk };
enum struct CeresSolverType { LinearSearch_BFGS, LinearSearch_LBFGS, TrustRegion_DenseQR };

/* ------------------------------------------- parameters (types.h:93-316) */
struct ImagerPars {
    ImagerPars()
        : image_size(0), padded_image_size(0), cell_size(0.0), padding_factor(1.0),
          kernel_function(KernelFunction::PSWF), kernel_support(0), kernel_exact(true), oversampling(0),
          generate_beam(false), gridding_correction(false), analytic_gcf(false), r_fft(FFTRoutine::FFTW_ESTIMATE_FFT)
    {
    }
    ImagerPars(unsigned _image_size, double _cell_size, double _padding_factor = 1.0,
        KernelFunction _kernel_function = KernelFunction::PSWF, unsigned _kernel_support = 3, bool _kernel_exact = true,
        unsigned _oversampling = 1, bool _generate_beam = false, bool _gridding_correction = true, bool _analytic_gcf = true,
        FFTRoutine _r_fft = FFTRoutine::FFTW_ESTIMATE_FFT, const std::string& _fft_wisdom_filename = std::string())
        : image_size(_image_size), cell_size(_cell_size), padding_factor(_padding_factor), kernel_function(_kernel_function),
          kernel_support(_kernel_support), kernel_exact(_kernel_exact), oversampling(_oversampling),
          generate_beam(_generate_beam), gridding_correction(_gridding_correction), analytic_gcf(_analytic_gcf),
          r_fft(_r_fft), fft_wisdom_filename(_fft_wisdom_filename)
    {
        padded_image_size = unsigned(image_size * padding_factor);
    }
    unsigned image_size;
    unsigned padded_image_size;
    double cell_size;
    double padding_factor;
    KernelFunction kernel_function;
    unsigned kernel_support;
    bool kernel_exact;
    unsigned oversampling;
    bool generate_beam;
    bool gridding_correction;
    bool analytic_gcf;
    FFTRoutine r_fft;
    std::string fft_wisdom_filename;
};

struct W_ProjectionPars {
    W_ProjectionPars()
        : num_wplanes(0), max_wpconv_support(0), undersampling_opt(1), kernel_trunc_perc(0.0), hankel_opt(false),
          hankel_proj_slice(false), interp_type(InterpType::LINEAR), wplanes_median(false)
    {
    }
    W_ProjectionPars(unsigned _num_wplanes, unsigned _max_wpconv_support, unsigned _undersampling_opt = 1,
        double _kernel_trunc_perc = 0.0, bool _hankel_opt = false, bool _hankel_proj_slice = false,
        InterpType _interp_type = InterpType::LINEAR, bool _wplanes_median = false)
        : num_wplanes(_num_wplanes), max_wpconv_support(_max_wpconv_support), undersampling_opt(_undersampling_opt),
          kernel_trunc_perc(_kernel_trunc_perc), hankel_opt(_hankel_opt), hankel_proj_slice(_hankel_proj_slice),
          interp_type(_interp_type), wplanes_median(_wplanes_median)
    {
    }
    bool isEnabled() const { return num_wplanes > 0; }
    unsigned num_wplanes;
    unsigned max_wpconv_support;
    unsigned undersampling_opt;
    double kernel_trunc_perc;
    bool hankel_opt;
    bool hankel_proj_slice;
    InterpType interp_type;
    bool wplanes_median;
};

struct A_ProjectionPars {
    A_ProjectionPars()
        : num_timesteps(0), obs_dec(0.0), obs_ra(0.0), aproj_opt(false), aproj_mask_perc(0.0)
    {
    }
    A_ProjectionPars(unsigned _num_timesteps, double _obs_dec = 0.0, double _obs_ra = 0.0, bool _aproj_opt = false,
        double _aproj_mask_perc = 0.0, const mat& _lha = mat(), const std::vector<double>& _pbeam_coefs = {})
        : num_timesteps(_num_timesteps), obs_dec(_obs_dec), obs_ra(_obs_ra), aproj_opt(_aproj_opt),
          aproj_mask_perc(_aproj_mask_perc), lha(_lha), pbeam_coefs(_pbeam_coefs)
    {
    }
    bool isEnabled() const { return num_timesteps > 0; }
    unsigned num_timesteps;
    double obs_dec, obs_ra;
    bool aproj_opt;
    double aproj_mask_perc;
    mat lha;
    std::vector<double> pbeam_coefs;
};

/* -------------------------------------- kernel functors (conv_func.h:23-262) */
struct KernelFunctor {
    stp_kernel_spec spec;
};
struct TopHat : KernelFunctor {
    explicit TopHat(double half_base_width) { stp_kernel_spec_init(&spec, STP_KERNEL_TOPHAT, half_base_width); }
};
struct Triangle : KernelFunctor {
    explicit Triangle(double half_base_width) { stp_kernel_spec_init(&spec, STP_KERNEL_TRIANGLE, half_base_width); }
};
struct Sinc : KernelFunctor {
    Sinc(double trunc = 0.0, double width_normalization = 1.0)
    {
        stp_kernel_spec_init(&spec, STP_KERNEL_SINC, trunc);
        spec.width_normalization_sinc = width_normalization;
    }
};
struct Gaussian : KernelFunctor {
    Gaussian(double trunc = 0.0, double width_normalization = 1.0)
    {
        stp_kernel_spec_init(&spec, STP_KERNEL_GAUSSIAN, trunc);
        spec.width_normalization_gaussian = width_normalization;
    }
};
struct GaussianSinc : KernelFunctor {
    GaussianSinc(double trunc = 0.0, double width_normalization_gaussian = 2.52, double width_normalization_sinc = 1.55)
    {
        stp_kernel_spec_init(&spec, STP_KERNEL_GAUSSIANSINC, trunc);
        spec.width_normalization_gaussian = width_normalization_gaussian;
        spec.width_normalization_sinc = width_normalization_sinc;
    }
};
struct PSWF : KernelFunctor {
    explicit PSWF(double trunc = 0.0) { stp_kernel_spec_init(&spec, STP_KERNEL_PSWF, trunc); }
};

/* ------------------------------------------------------------------ context */
namespace detail {
struct Ctx {
    stp_ctx* h = nullptr;
    explicit Ctx(int device)
    {
        if (stp_ctx_create(&h, device) != STP_OK)
            throw std::runtime_error("stp: cannot create a context on HIP device " + std::to_string(device));
    }
    ~Ctx() { stp_ctx_destroy(h); }
};
inline int& device_slot()
{
    static thread_local int dev = 0;
    return dev;
}
inline stp_ctx* ctx()
{
    static thread_local std::unique_ptr<Ctx> c;
    static thread_local int dev = -1;
    if (!c || dev != device_slot()) {
        c.reset();
        c.reset(new Ctx(device_slot()));
        dev = device_slot();
    }
    return c->h;
}
inline void check(stp_status st)
{
    if (st != STP_OK)
        throw std::runtime_error(std::string("stp: ") + stp_ctx_last_error(ctx()));
}
inline stp_imager_pars to_c(const ImagerPars& p)
{
    stp_imager_pars c{};
    c.image_size = p.image_size;
    c.padded_image_size = p.padded_image_size;
    c.cell_size = p.cell_size;
    c.padding_factor = p.padding_factor;
    c.kernel_function = int32_t(p.kernel_function);
    c.kernel_support = p.kernel_support;
    c.kernel_exact = p.kernel_exact;
    c.oversampling = p.oversampling;
    c.generate_beam = p.generate_beam;
    c.gridding_correction = p.gridding_correction;
    c.analytic_gcf = p.analytic_gcf;
    c.r_fft = int32_t(p.r_fft);
    return c;
}
inline stp_wproj_pars to_c(const W_ProjectionPars& p)
{
    stp_wproj_pars c{};
    c.num_wplanes = p.num_wplanes;
    c.max_wpconv_support = p.max_wpconv_support;
    c.undersampling_opt = p.undersampling_opt;
    c.kernel_trunc_perc = p.kernel_trunc_perc;
    c.hankel_opt = p.hankel_opt;
    c.hankel_proj_slice = p.hankel_proj_slice;
    c.interp_type = int32_t(p.interp_type);
    c.wplanes_median = p.wplanes_median;
    return c;
}
inline stp_aproj_pars to_c(const A_ProjectionPars& p)
{
    stp_aproj_pars c{};
    c.num_timesteps = p.num_timesteps;
    c.obs_dec = p.obs_dec;
    c.obs_ra = p.obs_ra;
    c.aproj_opt = p.aproj_opt;
    c.aproj_mask_perc = p.aproj_mask_perc;
    return c;
}
inline std::pair<mat, mat> run_imager(const stp_kernel_spec* k, const cx_mat& vis, const mat& vis_weights,
    const mat& uvw_lambda, const ImagerPars& img_pars, const W_ProjectionPars& w_proj, const A_ProjectionPars& a_proj)
{
    if (uvw_lambda.n_cols != 3 || uvw_lambda.n_rows != vis.n_elem || vis.n_elem != vis_weights.n_elem)
        throw std::runtime_error("stp: vis, vis_weights and uvw_lambda (n x 3) sizes disagree");
    const stp_imager_pars ip = to_c(img_pars);
    const stp_wproj_pars wp = to_c(w_proj);
    const stp_aproj_pars ap = to_c(a_proj);
    mat image(img_pars.image_size, img_pars.image_size), beam;
    if (img_pars.generate_beam)
        beam.set_size(img_pars.image_size, img_pars.image_size);
    stp_imager_diag diag{};
    check(stp_image_visibilities(ctx(), &ip, &wp, &ap, k, vis.n_elem, reinterpret_cast<const double*>(vis.memptr()),
        vis_weights.memptr(), uvw_lambda.memptr(), STP_PRECISION_FP64, image.memptr(),
        img_pars.generate_beam ? beam.memptr() : nullptr, &diag));
    if (diag.image_rows == 0) { // sample_grid_total <= 0: empty result (imager.h:355)
        image.reset();
        beam.reset();
    }
    return { std::move(image), std::move(beam) };
}
} // namespace detail

/* Selects the HIP device used by the calling thread. */
inline void set_device(int device) { detail::device_slot() = device; }

/* ---------------------------------------------- imager (imager.h:233-411) */
template <typename T>
std::pair<mat, mat> image_visibilities(const T kernel_creator, const cx_mat& vis, const mat& vis_weights,
    const mat& uvw_lambda, const ImagerPars& img_pars = ImagerPars(), const W_ProjectionPars& w_proj = W_ProjectionPars(),
    const A_ProjectionPars& a_proj = A_ProjectionPars())
{
    return detail::run_imager(&kernel_creator.spec, vis, vis_weights, uvw_lambda, img_pars, w_proj, a_proj);
}

class ImageVisibilities {
public:
    ImageVisibilities() = delete;
    ImageVisibilities(const cx_mat& vis, const mat& vis_weights, const mat& uvw_lambda, const ImagerPars& img_pars,
        const W_ProjectionPars& w_proj = W_ProjectionPars(), const A_ProjectionPars& a_proj = A_ProjectionPars())
    {
        auto r = detail::run_imager(nullptr, vis, vis_weights, uvw_lambda, img_pars, w_proj, a_proj);
        vis_grid = std::move(r.first);
        sampling_grid = std::move(r.second);
    }
    mat vis_grid;
    mat sampling_grid;
};

/* ------------------------------ model visibilities (visibility.h:34-59) */
/* skymodel rows (l, m, flux_jy); uvw_baselines n x 3 (wavelengths). */
inline cx_mat generate_visibilities_from_local_skymodel(const mat& skymodel, const mat& uvw_baselines)
{
    if (skymodel.n_cols != 3 || uvw_baselines.n_cols != 3)
        throw std::runtime_error("stp: skymodel and uvw_baselines must have 3 columns");
    cx_mat out(uvw_baselines.n_rows, 1);
    detail::check(stp_generate_visibilities_from_local_skymodel(detail::ctx(), skymodel.n_rows, skymodel.memptr(),
        uvw_baselines.n_rows, uvw_baselines.memptr(), reinterpret_cast<double*>(out.memptr())));
    return out;
}
inline cx_mat visibilities_for_point_source(const mat& dist_uvw, double l, double m, double flux)
{
    mat sky(1, 3);
    sky.at(0, 0) = l;
    sky.at(0, 1) = m;
    sky.at(0, 2) = flux;
    return generate_visibilities_from_local_skymodel(sky, dist_uvw);
}

/* ---------------------------------------- source finder (sourcefind.h) */
struct BoundingBox {
    int top = 0, bottom = 0, left = 0, right = 0;
};
struct Gaussian2dParams {
    double amplitude = 0, x_centre = 0, y_centre = 0, semimajor = 0, semiminor = 0, theta = 0;
};
struct IslandParams {
    int label_idx = 0;
    double extremum_val = 0;
    int extremum_y_idx = 0;
    int extremum_x_idx = 0;
    int sign = 0;
    int num_samples = 0;
    BoundingBox bounding_box;
    Gaussian2dParams moments_fit;
    Gaussian2dParams leastsq_fit;
    std::string ceres_report;
};

class SourceFindImage {
public:
    SourceFindImage() = delete;
    SourceFindImage(const mat& input_data, double input_detection_n_sigma, double input_analysis_n_sigma,
        double input_rms_est = 0.0, bool find_negative_sources = true, int sigma_clip_iters = 5,
        MedianMethod median_method = MedianMethod::BINAPPROX, bool gaussian_fitting = false,
        bool ccl_4connectivity = false, bool generate_labelmap = true, int source_min_area = 5,
        CeresDiffMethod = CeresDiffMethod::AnalyticDiff_SingleResBlk,
        CeresSolverType = CeresSolverType::LinearSearch_LBFGS)
        : detection_n_sigma(input_detection_n_sigma)
        , analysis_n_sigma(input_analysis_n_sigma)
        , fit_gaussian(gaussian_fitting)
    {
        stp_sourcefind_pars p{};
        p.detection_n_sigma = input_detection_n_sigma;
        p.analysis_n_sigma = input_analysis_n_sigma;
        p.rms_est = input_rms_est;
        p.find_negative_sources = find_negative_sources;
        p.sigma_clip_iters = sigma_clip_iters;
        p.median_method = int32_t(median_method);
        p.gaussian_fitting = 0; // the fit runs below, on this call's image and label map
        p.ccl_4connectivity = ccl_4connectivity;
        p.generate_labelmap = generate_labelmap;
        p.source_min_area = source_min_area;
        label_map.set_size(input_data.n_rows, input_data.n_cols);
        std::vector<stp_island> isl(1024);
        size_t n = 0;
        stp_sourcefind_stats st{};
        stp_status s = stp_source_find(detail::ctx(), input_data.memptr(), int32_t(input_data.n_rows),
            int32_t(input_data.n_cols), &p, label_map.memptr(), isl.data(), isl.size(), &n, &st);
        if (s == STP_ERR_CAPACITY) {
            isl.resize(n);
            s = stp_source_find(detail::ctx(), input_data.memptr(), int32_t(input_data.n_rows), int32_t(input_data.n_cols),
                &p, label_map.memptr(), isl.data(), isl.size(), &n, &st);
        }
        detail::check(s);
        rms_est = st.rms_est;
        bg_level = st.bg_level;
        islands.reserve(n);
        for (size_t i = 0; i < n; i++) {
            IslandParams ip;
            ip.label_idx = isl[i].label_idx;
            ip.extremum_val = isl[i].extremum_val;
            ip.extremum_y_idx = isl[i].extremum_y_idx;
            ip.extremum_x_idx = isl[i].extremum_x_idx;
            ip.sign = isl[i].sign;
            ip.num_samples = isl[i].num_samples;
            ip.bounding_box = { isl[i].bounding_box.top, isl[i].bounding_box.bottom, isl[i].bounding_box.left,
                isl[i].bounding_box.right };
            const stp_gaussian2d& g = isl[i].moments_fit;
            ip.moments_fit = { g.amplitude, g.x_centre, g.y_centre, g.semimajor, g.semiminor, g.theta };
            islands.push_back(std::move(ip));
        }
        if (gaussian_fitting && n > 0) {
            // IslandParams::leastsq_fit_gaussian_2d for every island (sourcefind.cpp:281-289)
            std::vector<stp_fit_result> fits(n);
            detail::check(stp_fit_islands(detail::ctx(), input_data.memptr(), int32_t(input_data.n_rows),
                int32_t(input_data.n_cols), label_map.memptr(), isl.data(), n, fits.data()));
            for (size_t i = 0; i < n; i++) {
                const stp_gaussian2d& f = fits[i].leastsq_fit;
                islands[i].leastsq_fit = { f.amplitude, f.x_centre, f.y_centre, f.semimajor, f.semiminor, f.theta };
                const stp_bounding_box& b = fits[i].bounding_box;
                islands[i].bounding_box = { b.top, b.bottom, b.left, b.right };
                islands[i].ceres_report = fits[i].status == 2
                    ? "not solved (fewer than 3 samples)"
                    : "Levenberg-Marquardt, " + std::to_string(fits[i].iterations) + " iterations, " +
                        (fits[i].status == 0 ? "converged" : "iteration limit");
            }
        }
    }
    Mat<int> label_map;
    double detection_n_sigma;
    double analysis_n_sigma;
    double rms_est = 0.0;
    real_t bg_level = 0.0;
    std::vector<IslandParams> islands;
    bool fit_gaussian;
};

} // namespace stp

#endif /* STP_STP_H */
