"""Drop-in for the reference's Python module `stp_python`
(src/stp-python/stp_python.cpp:17-290): the same two wrapper functions, argument
names, defaults, enums and return shapes, executed on the MI355X through the
C-ABI (fastimaging_amd.stp.Context).

    import fastimaging_amd.stp_python as stp_python
    img, beam = stp_python.image_visibilities_wrapper(vis, snr_weights, uvw_lambda, 1024, 1.0, ...)
    islands = stp_python.source_find_wrapper(img, 50, 50)

Differences are errors, never silent fallbacks: Gaussian least-squares fitting
and A-projection raise (they are not implemented on the GPU), as does any
device failure.
"""
import enum
import threading

import numpy as np

from . import capi
from .stp import Context


class KernelFunction(enum.IntEnum):
    """stp::KernelFunction (src/stp/types.h:32-39)."""
    TopHat = capi.KERNEL_TOPHAT
    Triangle = capi.KERNEL_TRIANGLE
    Sinc = capi.KERNEL_SINC
    Gaussian = capi.KERNEL_GAUSSIAN
    GaussianSinc = capi.KERNEL_GAUSSIANSINC
    PSWF = capi.KERNEL_PSWF


class FFTRoutine(enum.IntEnum):
    """stp::FFTRoutine (src/stp/types.h:41-48); every value runs the same GPU FFT."""
    FFTW_ESTIMATE_FFT = 0
    FFTW_MEASURE_FFT = 1
    FFTW_PATIENT_FFT = 2
    FFTW_WISDOM_FFT = 3
    FFTW_WISDOM_INPLACE_FFT = 4


class InterpType(enum.IntEnum):
    """stp::InterpType (src/stp/types.h:50-55)."""
    LINEAR = 0
    CUBIC = 1
    COSINE = 2


class MedianMethod(enum.IntEnum):
    """stp::MedianMethod (src/stp/types.h:57-63)."""
    ZEROMEDIAN = capi.MEDIAN_ZEROMEDIAN
    BINMEDIAN = capi.MEDIAN_BINMEDIAN
    BINAPPROX = capi.MEDIAN_BINAPPROX
    NTHELEMENT = capi.MEDIAN_NTHELEMENT


class CeresDiffMethod(enum.IntEnum):
    """stp::CeresDiffMethod (src/stp/types.h:65-71)."""
    AutoDiff = 0
    AutoDiff_SingleResBlk = 1
    AnalyticDiff = 2
    AnalyticDiff_SingleResBlk = 3


class CeresSolverType(enum.IntEnum):
    """stp::CeresSolverType (src/stp/types.h:73-78)."""
    LinearSearch_BFGS = 0
    LinearSearch_LBFGS = 1
    TrustRegion_DenseQR = 2


class Gaussian2dParams:
    """stp::Gaussian2dParams (src/stp/sourcefind/fitting.h:60-152), the bound fields."""

    def __init__(self, amplitude=0.0, x_centre=0.0, y_centre=0.0, semimajor=0.0, semiminor=0.0, theta=0.0):
        self.amplitude = amplitude
        self.x_centre = x_centre
        self.y_centre = y_centre
        self.semimajor = semimajor
        self.semiminor = semiminor
        self.theta = theta

    def __repr__(self):
        # stp_python.cpp:231-235 (std::to_string prints 6 decimals)
        return "Gaussian2dParams=[" + ", ".join(f"{v:.6f}" for v in (
            self.amplitude, self.x_centre, self.y_centre, self.semimajor, self.semiminor, self.theta)) + "]"


_tls = threading.local()


def _ctx():
    c = getattr(_tls, "ctx", None)
    if c is None:
        c = Context(0)
        _tls.ctx = c
    return c


def image_visibilities_wrapper(vis, snr_weights, uvw_lambda, image_size, cell_size, padding_factor=1.0,
                               kernel_func=KernelFunction.PSWF, kernel_support=3, kernel_exact=True,
                               kernel_oversampling=8, generate_beam=False, gridding_correction=True,
                               analytic_gcf=False, r_fft=FFTRoutine.FFTW_ESTIMATE_FFT, fft_wisdom_filename="",
                               num_wplanes=0, wplanes_median=False, max_wpconv_support=0, hankel_opt=False,
                               hankel_proj_slice=False, undersampling_opt=1, kernel_trunc_perc=0.0,
                               interp_type=InterpType.LINEAR, aproj_numtimesteps=0, obs_dec=0.0, obs_ra=0.0,
                               aproj_opt=False, aproj_mask_perc=0.0, lha=None, pbeam_coefs=None):
    """image_visibilities_wrapper (stp_python.cpp:17-138): returns (image, beam) as
    N x N float64 arrays (the beam is empty when generate_beam is False)."""
    vis = np.asarray(vis)
    uvw = np.asarray(uvw_lambda, dtype=np.float64)
    if vis.ndim != 1 or uvw.ndim != 2:
        raise ValueError("vis must be 1-D and uvw_lambda 2-D (stp_python.cpp:49-50)")
    # image sizes are rounded up to multiples of 4 (stp_python.cpp:59-67)
    image_size = int(image_size)
    while image_size % 4:
        image_size += 1
    padded = int(float(image_size) * padding_factor)
    while padded % 4:
        padded += 1
    pars = capi.imager_pars(image_size, cell_size, kernel_function=int(kernel_func), kernel_support=kernel_support,
                            kernel_exact=kernel_exact, oversampling=kernel_oversampling, generate_beam=generate_beam,
                            gridding_correction=gridding_correction, analytic_gcf=analytic_gcf,
                            padding_factor=padding_factor)
    pars.padded_image_size = padded
    pars.r_fft = int(r_fft)
    wp = None
    if num_wplanes > 0:
        wp = capi.wproj_pars(num_wplanes, max_wpconv_support, undersampling_opt, kernel_trunc_perc, wplanes_median)
        wp.hankel_opt = int(hankel_opt)
        wp.hankel_proj_slice = int(hankel_proj_slice)
        wp.interp_type = int(interp_type)
    ap = None
    if aproj_numtimesteps > 0:
        ap = capi.AProjPars()
        ap.num_timesteps = aproj_numtimesteps
        ap.obs_dec = obs_dec
        ap.obs_ra = obs_ra
        ap.aproj_opt = int(aproj_opt)
        ap.aproj_mask_perc = aproj_mask_perc
    img, beam = _ctx().image_visibilities(vis, snr_weights, uvw, pars, w_proj=wp, a_proj=ap)
    if beam is None:
        beam = np.empty((0, 0), order="F")
    return img, beam


def source_find_wrapper(image_data, detection_n_sigma, analysis_n_sigma, rms_est=0.0, find_negative_sources=True,
                        sigma_clip_iters=5, median_method=MedianMethod.BINAPPROX, gaussian_fitting=False,
                        ccl_4connectivity=False, generate_labelmap=True, source_min_area=5,
                        ceres_diffmethod=CeresDiffMethod.AutoDiff_SingleResBlk,
                        ceres_solvertype=CeresSolverType.LinearSearch_BFGS):
    """source_find_wrapper (stp_python.cpp:140-185): returns a list of tuples
    (sign, extremum_val, extremum_x_idx, extremum_y_idx, num_samples, moments_fit,
    leastsq_fit, ceres_report)."""
    img = np.asarray(image_data)
    if img.ndim != 2:
        raise ValueError("image_data must be 2-D (stp_python.cpp:158)")
    pars = capi.sourcefind_pars(detection_n_sigma, analysis_n_sigma, rms_est=rms_est,
                                find_negative_sources=find_negative_sources, sigma_clip_iters=sigma_clip_iters,
                                median_method=int(median_method), gaussian_fitting=False,
                                ccl_4connectivity=ccl_4connectivity, generate_labelmap=generate_labelmap,
                                source_min_area=source_min_area)
    img64 = np.asfortranarray(img, dtype=np.float64)
    ctx = _ctx()
    res = ctx.source_find(img64, pars, label_map=True)
    # gaussian_fitting: the least-squares fit of every island (sourcefind.cpp:281-289)
    fits = ctx.fit_islands(img64, res.label_map, res.islands) if gaussian_fitting else None
    out = []
    for k, i in enumerate(res.islands):
        mf = Gaussian2dParams(*i["moments_fit"])
        lf, report = Gaussian2dParams(), ""
        if fits is not None:
            lf = Gaussian2dParams(*fits[k]["leastsq_fit"])
            report = (f"Levenberg-Marquardt: {fits[k]['iterations']} iterations, "
                      + ("not solved (fewer than 3 samples)" if fits[k]["status"] == 2 else
                         "converged" if fits[k]["status"] == 0 else "iteration limit"))
        out.append((i["sign"], i["extremum_val"], i["extremum_x_idx"], i["extremum_y_idx"], i["num_samples"], mf,
                    lf, report))
    return out
