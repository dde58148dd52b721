/*
 * stp_capi.h - C-ABI of the MI355X (gfx950) Slow Transients Pipeline imaging and
 * source-finding hot path.
 *
 * Plain C types only: POD structs, raw pointers and sizes, integer status codes.
 * No exception ever crosses this boundary; every entry point returns an
 * stp_status and leaves a human-readable message in stp_ctx_last_error().
 *
 * Matrix layout everywhere is column-major (Armadillo / Fortran order), the
 * layout of the reference's arma::Mat (src/stp/common/matstp.h:94):
 *   element (row r, col c) of an n_rows x n_cols matrix is at [r + c*n_rows].
 * Complex values are interleaved (re, im) doubles, as std::complex<double>.
 *
 * Each entry point names the reference interface it replaces (file:line, paths
 * relative to the reference repository root).
 */
#ifndef STP_CAPI_H
#define STP_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
typedef enum stp_status {
    STP_OK = 0,
    STP_ERR_INVALID_ARGUMENT = 1, /* reference: assert(...) in imager.h:257-287, gridder.h:270-295 */
    STP_ERR_UNSUPPORTED = 2,      /* feature outside this implementation (A-projection) */
    STP_ERR_DEVICE = 3,           /* HIP runtime error */
    STP_ERR_OUT_OF_MEMORY = 4,    /* device allocation failed or HBM budget exceeded */
    STP_ERR_INTERNAL = 5,         /* internal consistency check failed */
    STP_ERR_CAPACITY = 6          /* caller-provided output buffer too small */
} stp_status;

/* ------------------------------------------------------------------------- */
/* Enumerations (same order and meaning as src/stp/types.h:32-88)            */
/* ------------------------------------------------------------------------- */
typedef enum stp_kernel_function {
    STP_KERNEL_TOPHAT = 0,
    STP_KERNEL_TRIANGLE = 1,
    STP_KERNEL_SINC = 2,
    STP_KERNEL_GAUSSIAN = 3,
    STP_KERNEL_GAUSSIANSINC = 4,
    STP_KERNEL_PSWF = 5
} stp_kernel_function;

typedef enum stp_fft_routine {
    STP_FFT_ESTIMATE = 0,
    STP_FFT_MEASURE = 1,
    STP_FFT_PATIENT = 2,
    STP_FFT_WISDOM = 3,
    STP_FFT_WISDOM_INPLACE = 4
} stp_fft_routine;

typedef enum stp_median_method {
    STP_MEDIAN_ZEROMEDIAN = 0,
    STP_MEDIAN_BINMEDIAN = 1,
    STP_MEDIAN_BINAPPROX = 2,
    STP_MEDIAN_NTHELEMENT = 3
} stp_median_method;

/* Arithmetic of the gridder + FFT.  FP64 is the reference's USE_FLOAT=OFF build.
 * The FP32 forms keep the uv-to-pixel index and phase math, the kernel-sum
 * normalisation and the weight total in FP64, compute the taps in FP32, store
 * FP32 grids and images and run the FFT in FP32 (twiddles rounded from long
 * double).  They differ in how each grid cell's taps are summed in the LDS tile:
 *   FP32C           FP64 sums (ds_add_f64), rounded to FP32 once per cell;
 *   FP32_TWOFLOAT   two-float hi + lo sums: TwoSum on the old value returned by
 *                   the FP32 LDS atomic, the exact rounding error added to lo.
 * Source finding on an FP32 image reads FP32 and accumulates in FP64. */
typedef enum stp_precision {
    STP_PRECISION_FP64 = 0,
    STP_PRECISION_FP32C = 1,
    STP_PRECISION_FP32_TWOFLOAT = 2
} stp_precision;

/* ------------------------------------------------------------------------- */
/* Parameter structs                                                          */
/* ------------------------------------------------------------------------- */

/* stp::ImagerPars (src/stp/types.h:93-173).  padded_image_size is taken as given
 * (the reference computes image_size * padding_factor in its constructor,
 * types.h:157); stp_imager_pars_init() fills it the same way. */
typedef struct stp_imager_pars {
    uint32_t image_size;
    uint32_t padded_image_size;
    double cell_size;        /* arcsec */
    double padding_factor;
    int32_t kernel_function; /* stp_kernel_function */
    uint32_t kernel_support;
    int32_t kernel_exact;
    uint32_t oversampling;
    int32_t generate_beam;
    int32_t gridding_correction;
    int32_t analytic_gcf;
    int32_t r_fft;           /* stp_fft_routine; accepted and ignored (no FFTW) */
} stp_imager_pars;

/* Convolution-kernel functor (src/stp/convolution/conv_func.h:23-262):
 *   TopHat(half_base_width), Triangle(half_base_width), Sinc(trunc, width),
 *   Gaussian(trunc, width), GaussianSinc(trunc, width_gaussian, width_sinc),
 *   PSWF(trunc).  `param` is half_base_width or trunc; the widths apply to the
 *   functors that take them (Gaussian uses width_normalization_gaussian, Sinc
 *   width_normalization_sinc).  stp_kernel_spec_init() fills the reference's
 *   default widths (1.0, 1.0; GaussianSinc 2.52 and 1.55, conv_func.h:216-217). */
typedef struct stp_kernel_spec {
    int32_t kernel_function; /* stp_kernel_function */
    double param;
    double width_normalization_gaussian;
    double width_normalization_sinc;
} stp_kernel_spec;

/* stp::W_ProjectionPars (src/stp/types.h:178-249).  W-projection is on iff
 * num_wplanes > 0 (types.h:233-239). */
typedef struct stp_wproj_pars {
    uint32_t num_wplanes;
    uint32_t max_wpconv_support;
    uint32_t undersampling_opt;
    double kernel_trunc_perc;
    int32_t hankel_opt;        /* radially symmetric W-kernels by a Hankel transform (aw_projection.cpp:211-287) */
    int32_t hankel_proj_slice; /* with hankel_opt: the DFT of the kernel's projection instead of the DHT */
    int32_t interp_type;       /* with hankel_opt: 0 linear, 1 cosine (as linear), 2 cubic spline */
    int32_t wplanes_median;
} stp_wproj_pars;

/* stp::A_ProjectionPars (src/stp/types.h:254-316).  Accepted for signature
 * compatibility; num_timesteps > 0 returns STP_ERR_UNSUPPORTED. */
typedef struct stp_aproj_pars {
    uint32_t num_timesteps;
    double obs_dec;
    double obs_ra;
    int32_t aproj_opt;
    double aproj_mask_perc;
} stp_aproj_pars;

/* Arguments of the stp::SourceFindImage constructor (src/stp/sourcefind/sourcefind.h:188-201). */
typedef struct stp_sourcefind_pars {
    double detection_n_sigma;
    double analysis_n_sigma;
    double rms_est;              /* <= 0: estimate from the image */
    int32_t find_negative_sources;
    int32_t sigma_clip_iters;
    int32_t median_method;       /* stp_median_method */
    int32_t gaussian_fitting;    /* must be 0 here; the fit is stp_fit_islands */
    int32_t ccl_4connectivity;
    int32_t generate_labelmap;
    int32_t source_min_area;
} stp_sourcefind_pars;

/* Output record types (src/stp/sourcefind/fitting.h:18-152, sourcefind.h:63-133). */
typedef struct stp_bounding_box {
    int32_t top, bottom, left, right;
} stp_bounding_box;

typedef struct stp_gaussian2d {
    double amplitude, x_centre, y_centre, semimajor, semiminor, theta;
} stp_gaussian2d;

typedef struct stp_island {
    int32_t label_idx;        /* +id for positive islands, -id for negative ones */
    int32_t sign;
    int32_t extremum_y_idx;   /* visual (fft-shifted) coordinates, sourcefind.cpp:243-244 */
    int32_t extremum_x_idx;
    int32_t num_samples;
    int32_t reserved;
    double extremum_val;
    stp_bounding_box bounding_box;
    stp_gaussian2d moments_fit;
} stp_island;

/* Result of the least-squares Gaussian fit of one island
 * (IslandParams::leastsq_fit and bounding_box, sourcefind.h:95-133). */
typedef struct stp_fit_result {
    stp_gaussian2d leastsq_fit;     /* constrained as convert_to_constrained_parameters */
    stp_bounding_box bounding_box;  /* visual coordinates (sourcefind.cpp:449-666) */
    int32_t iterations;             /* linear solves of the minimiser (0: not solved) */
    int32_t status;                 /* 0 converged, 1 iteration limit, 2 fewer than 3 samples */
} stp_fit_result;

/* Convergence parameters of the island fit (Levenberg-Marquardt with Marquardt's
 * diagonal scaling and Nielsen's damping update; see stp_fit_islands). */
#define STP_FIT_MU0 1.0e-4        /* initial damping */
#define STP_FIT_MU_MAX 1.0e32     /* damping at which the search stops */
#define STP_FIT_DIAG_MIN 1.0e-6   /* clamp of the J^T J diagonal used for scaling */
#define STP_FIT_DIAG_MAX 1.0e32
#define STP_FIT_FTOL 1.0e-14      /* accepted step: relative decrease of the cost */
#define STP_FIT_XTOL 1.0e-12      /* accepted step: |step| <= XTOL (|x| + XTOL) */
#define STP_FIT_MAX_ITERS 200     /* linear solves per island */

/* Scalars the reference exposes as SourceFindImage members (sourcefind.h:146-164). */
typedef struct stp_sourcefind_stats {
    double bg_level;
    double rms_est;
    double mean;              /* from the one-pass mean/sigma (0 if not computed) */
    double sigma;
    uint32_t num_labels_pos;  /* labels produced by the CCL before detection filtering */
    uint32_t num_labels_neg;
    uint32_t num_valid_labels;
    uint32_t num_islands;     /* islands with area >= source_min_area */
} stp_sourcefind_stats;

/* Diagnostics of one imaging call. */
typedef struct stp_imager_diag {
    double sample_grid_total;   /* 2 * sum of good weights (gridder.h:415-423) */
    uint64_t num_good_vis;      /* visibilities passing the bounds check */
    uint32_t image_rows;        /* 0 when sample_grid_total <= 0 (empty result, imager.h:355) */
    uint32_t image_cols;
    double ms_total;            /* device time of the whole call (ms) */
    uint64_t num_grid_entries;  /* (visibility pass, grid tile) entries the gridder binned */
} stp_imager_diag;

/* ------------------------------------------------------------------------- */
/* Context                                                                    */
/* ------------------------------------------------------------------------- */
typedef struct stp_ctx stp_ctx;

/* Creates a context bound to HIP device `device` with its own stream and a
 * grow-only device workspace.  Replaces the reference's process-global FFTW
 * thread and wisdom setup (src/stp/common/fft.cpp:15-43). */
stp_status stp_ctx_create(stp_ctx** ctx, int device);
stp_status stp_ctx_destroy(stp_ctx* ctx);
const char* stp_ctx_last_error(const stp_ctx* ctx);
stp_status stp_ctx_synchronize(stp_ctx* ctx);
/* Releases every cached device buffer of the context (workspace, kernel tables,
 * FFT twiddles) after waiting for the context's own streams. */
stp_status stp_ctx_release_workspace(stp_ctx* ctx);
/* Bytes of device memory the context currently holds (workspace, kernel tables,
 * FFT twiddles). */
size_t stp_ctx_workspace_bytes(const stp_ctx* ctx);

/* Fills the defaults of the reference's ImagerPars full constructor
 * (types.h:132-158) for the given size and cell. */
void stp_imager_pars_init(stp_imager_pars* p, uint32_t image_size, double cell_size);
void stp_kernel_spec_init(stp_kernel_spec* k, int32_t kernel_function, double param);
void stp_wproj_pars_init(stp_wproj_pars* p);
void stp_sourcefind_pars_init(stp_sourcefind_pars* p, double detection_n_sigma, double analysis_n_sigma);

/* ------------------------------------------------------------------------- */
/* Imaging                                                                    */
/* ------------------------------------------------------------------------- */

/* stp::image_visibilities<T>(kernel_creator, vis, vis_weights, uvw_lambda,
 *                            img_pars, w_proj, a_proj)
 *   src/stp/imager/imager.h:233-376 (dispatch: src/stp/imager/imager.cpp:13-69)
 * kernel: the kernel_creator functor; NULL selects the functor the way the
 * ImageVisibilities constructor does (img_pars->kernel_function with
 * trunc / half_base_width = kernel_support, imager.cpp:25-61).
 * Host buffers: vis = n_vis complex (2*n_vis doubles), vis_weights = n_vis,
 * uvw_lambda = n_vis x 3 column-major (u column, v column, w column).
 * image and beam: image_size x image_size column-major, caller-allocated; beam
 * may be NULL when generate_beam == 0.  diag may be NULL.  Synchronous. */
stp_status stp_image_visibilities(stp_ctx* ctx,
    const stp_imager_pars* img_pars,
    const stp_wproj_pars* w_proj,   /* NULL = disabled */
    const stp_aproj_pars* a_proj,   /* NULL = disabled */
    const stp_kernel_spec* kernel,  /* NULL = dispatch on img_pars */
    size_t n_vis,
    const double* vis,
    const double* vis_weights,
    const double* uvw_lambda,
    int32_t precision,              /* stp_precision */
    double* image,
    double* beam,
    stp_imager_diag* diag);

/* Same computation with every buffer already in device memory of the context's
 * device.  Enqueued on the context stream; returns before the work finishes
 * unless `synchronize` is non-zero.  diag->sample_grid_total is only valid
 * after synchronisation.  Image/beam element type: double for FP64, float for
 * FP32C. */
stp_status stp_image_visibilities_device(stp_ctx* ctx,
    const stp_imager_pars* img_pars,
    const stp_wproj_pars* w_proj,
    const stp_aproj_pars* a_proj,
    const stp_kernel_spec* kernel,
    size_t n_vis,
    const double* d_vis,
    const double* d_vis_weights,
    const double* d_uvw_lambda,
    int32_t precision,
    void* d_image,
    void* d_beam,
    int32_t synchronize,
    stp_imager_diag* diag);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same computation, arguments and results as
 * stp_image_visibilities_device, with every buffer beyond the caller's inputs and
 * outputs (grids, gridding entries, kernel tables, FFT twiddles, the
 * sample-grid total) carved from one fixed device region of
 * STP_NOALLOC_SCRATCH_BYTES per device that the library holds until the process
 * ends.  The call makes no hipMalloc / hipFree / host-staging allocation.
 * A call whose buffers do not fit fails with STP_ERR_CAPACITY, leaving the outputs
 * unspecified (FP64 at N = 4096 needs 269 MB for its two half-plane grids alone and
 * is rejected).  No autotuning.  The region keeps the last call's contents; every buffer
 * that must start at zero is cleared by the call.  The call first waits for the
 * context stream; every call on one device shares the region, so calls that
 * overlap on one device (from other contexts or threads) race. */
#define STP_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
stp_status stp_image_visibilities_device_noalloc(stp_ctx* ctx,
    const stp_imager_pars* img_pars,
    const stp_wproj_pars* w_proj,
    const stp_aproj_pars* a_proj,
    const stp_kernel_spec* kernel,
    size_t n_vis,
    const double* d_vis,
    const double* d_vis_weights,
    const double* d_uvw_lambda,
    int32_t precision,
    void* d_image,
    void* d_beam,
    int32_t synchronize,
    stp_imager_diag* diag);

/* stp::convolve_to_grid<generateBeam>(kernel_creator, support, image_size, uv_lambda,
 *     vis, vis_weights, kernel_exact, oversampling, shift_uv, halfplane_gridding)
 *   src/stp/gridder/gridder.h:240-997 (no W-projection in this entry point).
 * uv_lambda: n_vis x 2 column-major, already in pixel units.  Outputs:
 * vis_grid and sampling_grid, rows x image_size complex column-major with
 * rows = image_size/2+1 if halfplane_gridding else image_size.  sampling_grid
 * may be NULL when generate_beam == 0. */
stp_status stp_convolve_to_grid(stp_ctx* ctx,
    const stp_kernel_spec* kernel,  /* the kernel_creator functor */
    uint32_t support,
    int32_t image_size,
    size_t n_vis,
    const double* uv_lambda,
    const double* vis,
    const double* vis_weights,
    int32_t kernel_exact,
    uint32_t oversampling,
    int32_t shift_uv,
    int32_t halfplane_gridding,
    int32_t generate_beam,
    double* vis_grid,
    double* sampling_grid,
    double* sample_grid_total);

/* stp::fft_fftw_c2r(input, output, r_fft)   src/stp/common/fft.cpp:45-135
 * Unnormalised inverse 2-D complex-to-real FFT of a half-plane grid of
 * in_rows = n/2+1 rows and n_cols columns (column-major complex) into an
 * n x n_cols real matrix, n = 2*(in_rows-1); n and n_cols powers of two, 4 to
 * 16384, n_cols <= n. */
stp_status stp_fft_c2r(stp_ctx* ctx, const double* in, int32_t in_rows, int32_t n_cols, double* out);

/* stp::fft_fftw_c2c(input, output, r_fft, forward)   src/stp/common/fft.cpp:227-323
 * Unnormalised 2-D complex FFT (sign -1 forward, +1 inverse) of a square
 * n x n column-major complex matrix. */
stp_status stp_fft_c2c(stp_ctx* ctx, const double* in, int32_t n, int32_t forward, double* out);

/* ------------------------------------------------------------------------- */
/* Model and residual visibilities                                            */
/* ------------------------------------------------------------------------- */

/* stp::generate_visibilities_from_local_skymodel(skymodel, uvw_baselines)
 *   src/stp/visibility/visibility.cpp:31-41 (per source: visibilities_for_point_source,
 *   visibility.cpp:15-29).
 * skymodel: n_src x 3 column-major, rows (l, m, flux_jy) with l^2 + m^2 <= 1;
 * uvw_lambda: n_vis x 3 column-major (wavelengths); model_vis: n_vis complex,
 * caller-allocated.  Synchronous. */
stp_status stp_generate_visibilities_from_local_skymodel(stp_ctx* ctx,
    size_t n_src,
    const double* skymodel,
    size_t n_vis,
    const double* uvw_lambda,
    double* model_vis);

/* The same with the baselines and visibilities in device memory (the sky model
 * stays in host memory).  subtract == 0: d_vis = model; subtract != 0:
 * d_vis -= model in place, the residual visibilities the reduce driver images
 * with its -d option (src/reduce/reduce.cpp:113-122).  Enqueued on the context
 * stream; returns before the work finishes unless `synchronize` is non-zero. */
stp_status stp_model_visibilities_device(stp_ctx* ctx,
    size_t n_src,
    const double* skymodel,
    size_t n_vis,
    const double* d_uvw_lambda,
    double* d_vis,
    int32_t subtract,
    int32_t synchronize);

/* ------------------------------------------------------------------------- */
/* Source finding                                                             */
/* ------------------------------------------------------------------------- */

/* stp::SourceFindImage(data, detection_n_sigma, analysis_n_sigma, rms_est,
 *     find_negative_sources, sigma_clip_iters, median_method, gaussian_fitting,
 *     ccl_4connectivity, generate_labelmap, source_min_area)
 *   src/stp/sourcefind/sourcefind.cpp:156-292, sourcefind.h:188-201
 * image: rows x cols column-major doubles (the un-shifted imager output).
 * label_map: rows x cols int32 (may be NULL).  islands: caller array of
 * capacity island_capacity; *n_islands receives the number found (if it
 * exceeds the capacity the call fails with STP_ERR_CAPACITY and *n_islands
 * holds the required count).  Synchronous. */
stp_status stp_source_find(stp_ctx* ctx,
    const double* image,
    int32_t rows,
    int32_t cols,
    const stp_sourcefind_pars* pars,
    int32_t* label_map,
    stp_island* islands,
    size_t island_capacity,
    size_t* n_islands,
    stp_sourcefind_stats* stats);

/* Same, with the image (and label map, if not NULL) in device memory.
 * image_precision selects the element type of d_image (double for FP64,
 * float for FP32C).  Islands and stats are returned in host memory; an island
 * buffer from stp_host_malloc is written by the device directly (no copy). */
stp_status stp_source_find_device(stp_ctx* ctx,
    const void* d_image,
    int32_t image_precision,
    int32_t rows,
    int32_t cols,
    const stp_sourcefind_pars* pars,
    int32_t* d_label_map,
    stp_island* islands,
    size_t island_capacity,
    size_t* n_islands,
    stp_sourcefind_stats* stats);

/* IslandParams::leastsq_fit_gaussian_2d (src/stp/sourcefind/sourcefind.cpp:719-866),
 * which SourceFindImage runs for every island when gaussian_fitting is set
 * (sourcefind.cpp:281-289), with the bounding boxes it then computes
 * (sourcefind.cpp:449-666).  Call it after stp_source_find[_device] with the same
 * image, that call's label map and its islands.  Each island's pixels are the
 * label map entries equal to its label_idx; the residuals are the reference's
 * model minus data (fitting.cpp:59-120), started from the moments fit
 * (non-positive axes start at 1, sourcefind.cpp:736-743); islands of fewer than
 * 3 samples get (extremum_val, 0, 0, 0, 0, 0) and status 2 (sourcefind.cpp:724-728).
 * The reference minimises with Ceres (LBFGS by default); this library runs
 * Levenberg-Marquardt on the GPU (one wavefront per island) to the convergence
 * parameters STP_FIT_*, i.e. to the least-squares minimum itself.  A label map
 * that disagrees with an island's num_samples fails with STP_ERR_INVALID_ARGUMENT.
 * results: n_islands records (host).  Synchronous. */
stp_status stp_fit_islands(stp_ctx* ctx,
    const double* image,            /* rows x cols, host */
    int32_t rows,
    int32_t cols,
    const int32_t* label_map,       /* rows x cols, host */
    const stp_island* islands,
    size_t n_islands,
    stp_fit_result* results);

/* The same with the image (element type by image_precision) and the label map
 * in device memory. */
stp_status stp_fit_islands_device(stp_ctx* ctx,
    const void* d_image,
    int32_t image_precision,
    int32_t rows,
    int32_t cols,
    const int32_t* d_label_map,
    const stp_island* islands,
    size_t n_islands,
    stp_fit_result* results);

/* ------------------------------------------------------------------------- */
/* Device memory helpers (so callers need no HIP headers)                     */
/* ------------------------------------------------------------------------- */
stp_status stp_device_malloc(stp_ctx* ctx, void** ptr, size_t bytes);
stp_status stp_device_free(stp_ctx* ctx, void* ptr);
stp_status stp_memcpy_htod(stp_ctx* ctx, void* dst, const void* src, size_t bytes);
stp_status stp_memcpy_dtoh(stp_ctx* ctx, void* dst, const void* src, size_t bytes);
/* No reference counterpart.  Page-locked host memory: an island list written by
 * stp_source_find / stp_source_find_device into such a buffer arrives by DMA,
 * with no staging copy on the host (ordinary host memory works too). */
stp_status stp_host_malloc(stp_ctx* ctx, void** ptr, size_t bytes);
stp_status stp_host_free(stp_ctx* ctx, void* ptr);

/* ------------------------------------------------------------------------- */
/* Autotuner and batches                                                      */
/* ------------------------------------------------------------------------- */
/* No reference counterpart.  With autotuning on, the first imaging call of a
 * shape (padded size, support, oversampling, kernel mode, precision, W planes,
 * visibility-count octave) times the gridder tile widths 16 / 32 / 64 columns
 * and the u-axis FFT splits next to the balanced one on that call's input and
 * keeps the fastest for the context; `tune_file` (NULL: none) is read at once
 * and appended with every new choice.  mode: 0 off, 1 tune new shapes. */
stp_status stp_ctx_set_autotune(stp_ctx* ctx, int32_t mode, const char* tune_file);
/* The configuration chosen for a shape (tile_cols 0: none yet). */
stp_status stp_ctx_get_tuning(stp_ctx* ctx, const stp_imager_pars* img_pars, const stp_wproj_pars* w_proj, size_t n_vis,
    int32_t precision, int32_t* tile_cols, int32_t* fft_l1, double* us);

/* One snapshot of a batch; every buffer is in device memory. */
typedef struct stp_batch_item {
    stp_imager_pars img_pars;
    stp_wproj_pars w_proj;       /* num_wplanes = 0: no W-projection */
    size_t n_vis;
    const double* d_vis;
    const double* d_vis_weights;
    const double* d_uvw_lambda;
    void* d_image;               /* image_size^2 of the precision's type */
    void* d_beam;                /* NULL unless generate_beam */
} stp_batch_item;

#define STP_BATCH_GRAPH 1u       /* capture the batch once, replay the hipGraph on later calls */
#define STP_BATCH_AUTOTUNE 2u    /* tune every new shape of the batch before running it */
#define STP_BATCH_FUSED 4u       /* one launch per stage over all eligible snapshots */

/* Images n_items snapshots.  They run on up to max_concurrent streams of the
 * context (0 = 4), each stream with its own workspace, so small snapshots fill
 * the device together.  With STP_BATCH_GRAPH the first call with a given batch
 * (parameters, sizes and buffer addresses) runs eagerly and records the whole
 * batch into one hipGraph; later calls with the same batch launch that graph
 * (the flag is for fixed, reused buffers: a context keeps the 16 most recent
 * graphs).  With STP_BATCH_FUSED the snapshots without W-projection whose
 * kernels are oversampled with at most 16 taps (and at most 24576 grid tiles and
 * 16.7 M visibilities) are imaged together, eagerly and with no capture, in up
 * to four chunks of about equal work: in each chunk one binning, one column scan,
 * one scatter and one tile launch per kernel and tile width class span all of its
 * tiles, one launch all of its totals, and one FFT per group of equal transform
 * shape all of the group's grids, so any new batch runs in a few dozen launches.
 * The chunks' binning runs first, then their tile sums on one stream beside their
 * FFTs on a second (chunk c's tile sums overlap chunk c-1's FFTs).  The other
 * snapshots take the streams (and the graph) as above, beside them.  The fused snapshots' entry, count and grid
 * buffers are held at once (worst-case entries, as for single calls).
 * sample_grid_totals (host, n_items, or NULL) receives 2 * sum of good weights
 * per snapshot; the call returns after the batch has finished. */
stp_status stp_image_batch(stp_ctx* ctx, size_t n_items, const stp_batch_item* items, const stp_kernel_spec* kernel,
    int32_t precision, uint32_t max_concurrent, uint32_t flags, double* sample_grid_totals);

/* ------------------------------------------------------------------------- */
/* Serving: a request loop with continuous batching                           */
/* ------------------------------------------------------------------------- */
/* No reference counterpart: the reference runs one snapshot per process call
 * (src/reduce/reduce.cpp:113-143).  A server owns `num_slots` contexts on one
 * device, each with its own stream and resident kernel tables, FFT twiddles and
 * grid workspace.  Submitted jobs queue in arrival order; a free slot takes the
 * next job as soon as the job's estimated device footprint fits in what is left
 * of `hbm_budget_bytes`, so jobs that arrive while others run are admitted as
 * slots and memory free up, and their kernels overlap on the device.
 * Host buffers of a job must stay valid until stp_server_wait returns for it. */
typedef struct stp_server stp_server;

typedef struct stp_server_cfg {
    int32_t device;
    uint32_t num_slots;         /* concurrent jobs (>= 1) */
    uint64_t hbm_budget_bytes;  /* admission limit; 0 = 80 % of the device's free memory */
} stp_server_cfg;

typedef struct stp_job {
    stp_imager_pars img_pars;
    stp_wproj_pars w_proj;      /* num_wplanes = 0: no W-projection */
    int32_t use_kernel_spec;    /* 0: kernel selected from img_pars like the ImageVisibilities ctor */
    stp_kernel_spec kernel;
    int32_t precision;          /* stp_precision */
    size_t n_vis;
    const double* vis;          /* host: n complex (re, im) */
    const double* vis_weights;  /* host: n */
    const double* uvw_lambda;   /* host: n x 3 column-major */
    double* image;              /* host out N x N, or NULL */
    double* beam;               /* host out N x N (generate_beam), or NULL */
    int32_t source_find;        /* run SourceFindImage on the image */
    stp_sourcefind_pars sf_pars;
    int32_t* label_map;         /* host out N x N, or NULL */
    stp_island* islands;        /* host out, capacity island_cap */
    size_t island_cap;
} stp_job;

typedef struct stp_job_result {
    int32_t status;             /* stp_status of the job */
    int32_t slot;               /* slot that ran it */
    size_t n_islands;
    stp_sourcefind_stats stats;
    stp_imager_diag diag;
    double t_submit, t_start, t_end;   /* seconds on the server's steady clock */
    uint64_t job_bytes;          /* what admission reserved for the job (stp_job_device_bytes) */
    uint64_t slot_bytes_before;  /* device bytes the slot held when it took the job */
    uint64_t slot_bytes_after;   /* ... and after it: at most slot_bytes_before + job_bytes */
    char error[256];
} stp_job_result;

stp_status stp_server_create(stp_server** server, const stp_server_cfg* cfg);
/* Waits for every submitted job, then frees the slots. */
stp_status stp_server_destroy(stp_server* server);
/* Queues a copy of *job; *job_id identifies it for poll / wait. */
stp_status stp_server_submit(stp_server* server, const stp_job* job, uint64_t* job_id);
/* *done = 1 once the job has finished (successfully or not). */
stp_status stp_server_poll(stp_server* server, uint64_t job_id, int32_t* done);
/* Blocks until the job has finished and fills *result; each id can be waited once. */
stp_status stp_server_wait(stp_server* server, uint64_t job_id, stp_job_result* result);
/* Estimated device bytes of a job (what admission reserves). */
uint64_t stp_job_device_bytes(const stp_job* job);
/* Jobs queued, running, and bytes reserved right now: the estimates of the running
 * jobs plus the device memory every slot keeps between jobs.  A job that does not
 * fit makes the idle slots release what they keep. */
stp_status stp_server_load(stp_server* server, uint32_t* queued, uint32_t* running, uint64_t* reserved_bytes);
const char* stp_server_last_error(const stp_server* server);

/* ------------------------------------------------------------------------- */
/* Kernel timing (HIP events on the context stream)                           */
/* ------------------------------------------------------------------------- */
/* When enabled, every kernel launch of the context is bracketed by HIP events
 * recorded on the context stream, and durations accumulate per kernel name. */
stp_status stp_ctx_set_timing(stp_ctx* ctx, int32_t enable);
/* Restricts the timing to the kernels named in `names` (comma-separated, e.g.
 * "grid_tiles,fft_u1"); NULL or "" times every launch. */
stp_status stp_ctx_set_timing_filter(stp_ctx* ctx, const char* names);
stp_status stp_ctx_reset_timings(stp_ctx* ctx);
/* Copies up to `cap` entries: names (NUL-terminated, 64 bytes each), total
 * milliseconds and launch counts.  *n receives the number of distinct kernels. */
stp_status stp_ctx_get_timings(stp_ctx* ctx, char* names, double* total_ms, uint64_t* counts, size_t cap, size_t* n);

#ifdef __cplusplus
}
#endif

#endif /* STP_CAPI_H */
