"""ctypes mirror of include/stp_capi.h and the loader of the HIP library.

The structs below must match include/stp_capi.h field for field.  The library
is loaded from fastimaging_amd/lib/libstp_hip.so (built in-tree); a missing
library raises instead of falling back to any other implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libstp_hip.so")

# enums (src/stp/types.h:32-88)
KERNEL_TOPHAT, KERNEL_TRIANGLE, KERNEL_SINC, KERNEL_GAUSSIAN, KERNEL_GAUSSIANSINC, KERNEL_PSWF = range(6)
MEDIAN_ZEROMEDIAN, MEDIAN_BINMEDIAN, MEDIAN_BINAPPROX, MEDIAN_NTHELEMENT = range(4)
PRECISION_FP64, PRECISION_FP32C, PRECISION_FP32_TWOFLOAT = 0, 1, 2
# size of the device scratch region of stp_image_visibilities_device_noalloc
NOALLOC_SCRATCH_BYTES = 128 << 20

STATUS_NAMES = {
    0: "STP_OK", 1: "STP_ERR_INVALID_ARGUMENT", 2: "STP_ERR_UNSUPPORTED", 3: "STP_ERR_DEVICE",
    4: "STP_ERR_OUT_OF_MEMORY", 5: "STP_ERR_INTERNAL", 6: "STP_ERR_CAPACITY",
}
STATUS_CODES = {v: k for k, v in STATUS_NAMES.items()}


class ImagerPars(ctypes.Structure):
    """stp::ImagerPars (src/stp/types.h:93-173)."""
    _fields_ = [
        ("image_size", ctypes.c_uint32),
        ("padded_image_size", ctypes.c_uint32),
        ("cell_size", ctypes.c_double),
        ("padding_factor", ctypes.c_double),
        ("kernel_function", ctypes.c_int32),
        ("kernel_support", ctypes.c_uint32),
        ("kernel_exact", ctypes.c_int32),
        ("oversampling", ctypes.c_uint32),
        ("generate_beam", ctypes.c_int32),
        ("gridding_correction", ctypes.c_int32),
        ("analytic_gcf", ctypes.c_int32),
        ("r_fft", ctypes.c_int32),
    ]


class KernelSpec(ctypes.Structure):
    """The kernel_creator functor of image_visibilities / convolve_to_grid
    (src/stp/convolution/conv_func.h:23-262)."""
    _fields_ = [
        ("kernel_function", ctypes.c_int32),
        ("param", ctypes.c_double),
        ("width_normalization_gaussian", ctypes.c_double),
        ("width_normalization_sinc", ctypes.c_double),
    ]


class WProjPars(ctypes.Structure):
    """stp::W_ProjectionPars (src/stp/types.h:178-249)."""
    _fields_ = [
        ("num_wplanes", ctypes.c_uint32),
        ("max_wpconv_support", ctypes.c_uint32),
        ("undersampling_opt", ctypes.c_uint32),
        ("kernel_trunc_perc", ctypes.c_double),
        ("hankel_opt", ctypes.c_int32),
        ("hankel_proj_slice", ctypes.c_int32),
        ("interp_type", ctypes.c_int32),
        ("wplanes_median", ctypes.c_int32),
    ]


class AProjPars(ctypes.Structure):
    _fields_ = [
        ("num_timesteps", ctypes.c_uint32),
        ("obs_dec", ctypes.c_double),
        ("obs_ra", ctypes.c_double),
        ("aproj_opt", ctypes.c_int32),
        ("aproj_mask_perc", ctypes.c_double),
    ]


class SourceFindPars(ctypes.Structure):
    """Arguments of stp::SourceFindImage (src/stp/sourcefind/sourcefind.h:188-201)."""
    _fields_ = [
        ("detection_n_sigma", ctypes.c_double),
        ("analysis_n_sigma", ctypes.c_double),
        ("rms_est", ctypes.c_double),
        ("find_negative_sources", ctypes.c_int32),
        ("sigma_clip_iters", ctypes.c_int32),
        ("median_method", ctypes.c_int32),
        ("gaussian_fitting", ctypes.c_int32),
        ("ccl_4connectivity", ctypes.c_int32),
        ("generate_labelmap", ctypes.c_int32),
        ("source_min_area", ctypes.c_int32),
    ]


class BoundingBox(ctypes.Structure):
    _fields_ = [("top", ctypes.c_int32), ("bottom", ctypes.c_int32), ("left", ctypes.c_int32), ("right", ctypes.c_int32)]


class Gaussian2d(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("amplitude", "x_centre", "y_centre", "semimajor", "semiminor", "theta")]


class Island(ctypes.Structure):
    """stp::IslandParams (src/stp/sourcefind/sourcefind.h:63-133)."""
    _fields_ = [
        ("label_idx", ctypes.c_int32),
        ("sign", ctypes.c_int32),
        ("extremum_y_idx", ctypes.c_int32),
        ("extremum_x_idx", ctypes.c_int32),
        ("num_samples", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("extremum_val", ctypes.c_double),
        ("bounding_box", BoundingBox),
        ("moments_fit", Gaussian2d),
    ]


class SourceFindStats(ctypes.Structure):
    _fields_ = [
        ("bg_level", ctypes.c_double),
        ("rms_est", ctypes.c_double),
        ("mean", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("num_labels_pos", ctypes.c_uint32),
        ("num_labels_neg", ctypes.c_uint32),
        ("num_valid_labels", ctypes.c_uint32),
        ("num_islands", ctypes.c_uint32),
    ]


class ImagerDiag(ctypes.Structure):
    _fields_ = [
        ("sample_grid_total", ctypes.c_double),
        ("num_good_vis", ctypes.c_uint64),
        ("image_rows", ctypes.c_uint32),
        ("image_cols", ctypes.c_uint32),
        ("ms_total", ctypes.c_double),
        ("num_grid_entries", ctypes.c_uint64),
    ]


class BatchItem(ctypes.Structure):
    """stp_batch_item: one snapshot of a batch, device buffers."""
    _fields_ = [
        ("img_pars", ImagerPars),
        ("w_proj", WProjPars),
        ("n_vis", ctypes.c_size_t),
        ("d_vis", ctypes.c_void_p),
        ("d_vis_weights", ctypes.c_void_p),
        ("d_uvw_lambda", ctypes.c_void_p),
        ("d_image", ctypes.c_void_p),
        ("d_beam", ctypes.c_void_p),
    ]


BATCH_GRAPH, BATCH_AUTOTUNE, BATCH_FUSED = 1, 2, 4


class ServerCfg(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("num_slots", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64)]


class Job(ctypes.Structure):
    """stp_job: one snapshot imaging (+ source finding) request."""
    _fields_ = [
        ("img_pars", ImagerPars),
        ("w_proj", WProjPars),
        ("use_kernel_spec", ctypes.c_int32),
        ("kernel", KernelSpec),
        ("precision", ctypes.c_int32),
        ("n_vis", ctypes.c_size_t),
        ("vis", ctypes.c_void_p),
        ("vis_weights", ctypes.c_void_p),
        ("uvw_lambda", ctypes.c_void_p),
        ("image", ctypes.c_void_p),
        ("beam", ctypes.c_void_p),
        ("source_find", ctypes.c_int32),
        ("sf_pars", SourceFindPars),
        ("label_map", ctypes.c_void_p),
        ("islands", ctypes.c_void_p),
        ("island_cap", ctypes.c_size_t),
    ]


class JobResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("slot", ctypes.c_int32),
        ("n_islands", ctypes.c_size_t),
        ("stats", SourceFindStats),
        ("diag", ImagerDiag),
        ("t_submit", ctypes.c_double),
        ("t_start", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("job_bytes", ctypes.c_uint64),
        ("slot_bytes_before", ctypes.c_uint64),
        ("slot_bytes_after", ctypes.c_uint64),
        ("error", ctypes.c_char * 256),
    ]


def imager_pars(image_size, cell_size, kernel_function=KERNEL_GAUSSIANSINC, kernel_support=3,
                kernel_exact=False, oversampling=9, generate_beam=True, gridding_correction=False,
                analytic_gcf=False, padding_factor=1.0):
    p = ImagerPars()
    p.image_size = image_size
    p.padding_factor = padding_factor
    p.padded_image_size = int(image_size * padding_factor)
    p.cell_size = cell_size
    p.kernel_function = kernel_function
    p.kernel_support = kernel_support
    p.kernel_exact = int(kernel_exact)
    p.oversampling = oversampling
    p.generate_beam = int(generate_beam)
    p.gridding_correction = int(gridding_correction)
    p.analytic_gcf = int(analytic_gcf)
    p.r_fft = 0
    return p


def kernel_spec(kernel_function, param, width_normalization_gaussian=None, width_normalization_sinc=None):
    """Functor spec with the reference's default widths (conv_func.h:106-170;
    GaussianSinc 2.52 / 1.55, conv_func.h:216-217) unless given."""
    if isinstance(kernel_function, KernelSpec):
        return kernel_function
    gs = kernel_function == KERNEL_GAUSSIANSINC
    k = KernelSpec()
    k.kernel_function = kernel_function
    k.param = param
    k.width_normalization_gaussian = (2.52 if gs else 1.0) if width_normalization_gaussian is None \
        else width_normalization_gaussian
    k.width_normalization_sinc = (1.55 if gs else 1.0) if width_normalization_sinc is None \
        else width_normalization_sinc
    return k


def wproj_pars(num_wplanes=0, max_wpconv_support=0, undersampling_opt=1, kernel_trunc_perc=0.0,
               wplanes_median=False, hankel_opt=False, hankel_proj_slice=False, interp_type=0):
    p = WProjPars()
    p.num_wplanes = num_wplanes
    p.max_wpconv_support = max_wpconv_support
    p.undersampling_opt = undersampling_opt
    p.kernel_trunc_perc = kernel_trunc_perc
    p.hankel_opt = int(hankel_opt)
    p.hankel_proj_slice = int(hankel_proj_slice)
    p.interp_type = int(interp_type)
    p.wplanes_median = int(wplanes_median)
    return p


def sourcefind_pars(detection_n_sigma, analysis_n_sigma, rms_est=0.0, find_negative_sources=True,
                    sigma_clip_iters=5, median_method=MEDIAN_BINAPPROX, gaussian_fitting=False,
                    ccl_4connectivity=False, generate_labelmap=True, source_min_area=5):
    """Defaults of the stp::SourceFindImage constructor (sourcefind.h:188-201)."""
    p = SourceFindPars()
    p.detection_n_sigma = detection_n_sigma
    p.analysis_n_sigma = analysis_n_sigma
    p.rms_est = rms_est
    p.find_negative_sources = int(find_negative_sources)
    p.sigma_clip_iters = sigma_clip_iters
    p.median_method = median_method
    p.gaussian_fitting = int(gaussian_fitting)
    p.ccl_4connectivity = int(ccl_4connectivity)
    p.generate_labelmap = int(generate_labelmap)
    p.source_min_area = source_min_area
    return p


class FitResult(ctypes.Structure):
    """Least-squares fit of one island (IslandParams::leastsq_fit, bounding_box)."""
    _fields_ = [("leastsq_fit", Gaussian2d), ("bounding_box", BoundingBox), ("iterations", ctypes.c_int32),
                ("status", ctypes.c_int32)]


def fit_tuple(f):
    """Plain-Python view of a FitResult record."""
    g, b = f.leastsq_fit, f.bounding_box
    return dict(leastsq_fit=(g.amplitude, g.x_centre, g.y_centre, g.semimajor, g.semiminor, g.theta),
                bounding_box=(b.top, b.bottom, b.left, b.right), iterations=f.iterations, status=f.status)


def islands_to_records(islands):
    """A ctypes Island array from island dicts (island_tuple) or Island records."""
    islands = list(islands)
    arr = (Island * max(1, len(islands)))()
    for k, isl in enumerate(islands):
        if isinstance(isl, Island):
            arr[k] = isl
            continue
        r = arr[k]
        r.label_idx, r.sign = isl["label_idx"], isl["sign"]
        r.extremum_y_idx, r.extremum_x_idx = isl["extremum_y_idx"], isl["extremum_x_idx"]
        r.num_samples, r.extremum_val = isl["num_samples"], isl["extremum_val"]
        (r.moments_fit.amplitude, r.moments_fit.x_centre, r.moments_fit.y_centre, r.moments_fit.semimajor,
         r.moments_fit.semiminor, r.moments_fit.theta) = isl["moments_fit"]
    return arr


def island_tuple(isl):
    """Plain-Python view of an Island record."""
    g = isl.moments_fit
    return dict(label_idx=isl.label_idx, sign=isl.sign, extremum_y_idx=isl.extremum_y_idx,
                extremum_x_idx=isl.extremum_x_idx, num_samples=isl.num_samples,
                extremum_val=isl.extremum_val,
                moments_fit=(g.amplitude, g.x_centre, g.y_centre, g.semimajor, g.semiminor, g.theta))


_lib = None

# exported symbols of the HIP library, with (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_SIG = {
    "stp_ctx_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "stp_ctx_destroy": (ctypes.c_int, [_P]),
    "stp_ctx_last_error": (ctypes.c_char_p, [_P]),
    "stp_ctx_synchronize": (ctypes.c_int, [_P]),
    "stp_ctx_release_workspace": (ctypes.c_int, [_P]),
    "stp_ctx_workspace_bytes": (ctypes.c_size_t, [_P]),
    "stp_imager_pars_init": (None, [ctypes.POINTER(ImagerPars), ctypes.c_uint32, ctypes.c_double]),
    "stp_kernel_spec_init": (None, [ctypes.POINTER(KernelSpec), ctypes.c_int32, ctypes.c_double]),
    "stp_wproj_pars_init": (None, [ctypes.POINTER(WProjPars)]),
    "stp_sourcefind_pars_init": (None, [ctypes.POINTER(SourceFindPars), ctypes.c_double, ctypes.c_double]),
    "stp_image_visibilities": (ctypes.c_int, [_P, ctypes.POINTER(ImagerPars), ctypes.POINTER(WProjPars),
                                              ctypes.POINTER(AProjPars), ctypes.POINTER(KernelSpec),
                                              ctypes.c_size_t, _P, _P, _P,
                                              ctypes.c_int32, _P, _P, ctypes.POINTER(ImagerDiag)]),
    "stp_image_visibilities_device": (ctypes.c_int, [_P, ctypes.POINTER(ImagerPars), ctypes.POINTER(WProjPars),
                                                     ctypes.POINTER(AProjPars), ctypes.POINTER(KernelSpec),
                                                     ctypes.c_size_t, _P, _P, _P,
                                                     ctypes.c_int32, _P, _P, ctypes.c_int32,
                                                     ctypes.POINTER(ImagerDiag)]),
    "stp_image_visibilities_device_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(ImagerPars), ctypes.POINTER(WProjPars),
                                                             ctypes.POINTER(AProjPars), ctypes.POINTER(KernelSpec),
                                                             ctypes.c_size_t, _P, _P, _P,
                                                             ctypes.c_int32, _P, _P, ctypes.c_int32,
                                                             ctypes.POINTER(ImagerDiag)]),
    "stp_fit_islands": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, ctypes.c_size_t, _P]),
    "stp_fit_islands_device": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P,
                                              ctypes.c_size_t, _P]),
    "stp_convolve_to_grid": (ctypes.c_int, [_P, ctypes.POINTER(KernelSpec), ctypes.c_uint32, ctypes.c_int32,
                                            ctypes.c_size_t, _P, _P, _P, ctypes.c_int32, ctypes.c_uint32,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P, _D]),
    "stp_fft_c2r": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, _P]),
    "stp_fft_c2c": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, _P]),
    "stp_generate_visibilities_from_local_skymodel": (ctypes.c_int, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P]),
    "stp_model_visibilities_device": (ctypes.c_int, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, ctypes.c_int32,
                                                     ctypes.c_int32]),
    "stp_source_find": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(SourceFindPars), _P,
                                       ctypes.POINTER(Island), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(SourceFindStats)]),
    "stp_source_find_device": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.POINTER(SourceFindPars), _P, ctypes.POINTER(Island),
                                              ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                              ctypes.POINTER(SourceFindStats)]),
    "stp_device_malloc": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_size_t]),
    "stp_device_free": (ctypes.c_int, [_P, _P]),
    "stp_host_malloc": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_size_t]),
    "stp_host_free": (ctypes.c_int, [_P, _P]),
    "stp_memcpy_htod": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "stp_memcpy_dtoh": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "stp_ctx_set_autotune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p]),
    "stp_ctx_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ImagerPars), ctypes.POINTER(WProjPars), ctypes.c_size_t,
                                          ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                          _D]),
    "stp_image_batch": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(BatchItem), ctypes.POINTER(KernelSpec),
                                       ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, _D]),
    "stp_server_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(ServerCfg)]),
    "stp_server_destroy": (ctypes.c_int, [_P]),
    "stp_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]),
    "stp_server_poll": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32)]),
    "stp_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(JobResult)]),
    "stp_job_device_bytes": (ctypes.c_uint64, [ctypes.POINTER(Job)]),
    "stp_server_load": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint64)]),
    "stp_server_last_error": (ctypes.c_char_p, [_P]),
    "stp_ctx_set_timing": (ctypes.c_int, [_P, ctypes.c_int32]),
    "stp_ctx_set_timing_filter": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "stp_ctx_reset_timings": (ctypes.c_int, [_P]),
    "stp_ctx_get_timings": (ctypes.c_int, [_P, ctypes.c_char_p, _D, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
}

EXPORTED_SYMBOLS = tuple(_SIG)


def load(path=None):
    """Loads the HIP library (raises OSError if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (res, args) in _SIG.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib
