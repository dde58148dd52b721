"""JSON configuration and NPZ data of the reference's drivers.

ConfigurationFile mirrors src/auxiliary/load_json_config.h:29-197 (member names,
defaults and the string tables of load_json_config.cpp:17-140): sections
"imager_settings", "wprojection_settings", "aprojection_settings" and
"sourcefind_settings", each key optional.  load_npy_complex_array /
load_npy_double_array mirror src/auxiliary/load_data.h:19-149: a .npy file or one
array of an .npz, returned as a column-major matrix of the array's shape (1-D
arrays become one column).  Files are read with numpy's non-executing loader
(allow_pickle=False).
"""
import json

import numpy as np

from . import capi

KERNEL_FUNCTIONS = {"TopHat": capi.KERNEL_TOPHAT, "Triangle": capi.KERNEL_TRIANGLE, "Sinc": capi.KERNEL_SINC,
                    "Gaussian": capi.KERNEL_GAUSSIAN, "GaussianSinc": capi.KERNEL_GAUSSIANSINC,
                    "PSWF": capi.KERNEL_PSWF}
FFT_ROUTINES = {"FFTW_ESTIMATE_FFT": 0, "FFTW_MEASURE_FFT": 1, "FFTW_PATIENT_FFT": 2, "FFTW_WISDOM_FFT": 3,
                "FFTW_WISDOM_INPLACE_FFT": 4}
INTERP_TYPES = {"linear": 0, "cosine": 1, "cubic": 2}
MEDIAN_METHODS = {"ZEROMEDIAN": capi.MEDIAN_ZEROMEDIAN, "BINMEDIAN": capi.MEDIAN_BINMEDIAN,
                  "BINAPPROX": capi.MEDIAN_BINAPPROX, "NTHELEMENT": capi.MEDIAN_NTHELEMENT}
CERES_DIFF_METHODS = ("AutoDiff", "AutoDiff_SingleResBlk", "AnalyticDiff", "AnalyticDiff_SingleResBlk")
CERES_SOLVER_TYPES = ("LinearSearch_BFGS", "LinearSearch_LBFGS", "TrustRegion_DenseQR")


def _lookup(table, value, what):
    # the reference asserts on an unknown string (load_json_config.cpp:35,57,76,95,114,131)
    if value not in table:
        raise ValueError(f"unknown {what}: {value!r}")
    return table[value] if isinstance(table, dict) else value


class ConfigurationFile:
    """ConfigurationFile (load_json_config.h:29-197)."""

    def __init__(self, path=None, document=None):
        # the default-constructed ImagerPars (types.h:95-110)
        p = self.img_pars = capi.ImagerPars()
        p.image_size = p.padded_image_size = 0
        p.cell_size = 0.0
        p.padding_factor = 1.0
        p.kernel_function = capi.KERNEL_PSWF
        p.kernel_support = 0
        p.kernel_exact = 1
        p.oversampling = 0
        p.generate_beam = p.gridding_correction = p.analytic_gcf = 0
        p.r_fft = FFT_ROUTINES["FFTW_ESTIMATE_FFT"]
        self.fft_wisdom_filename = ""
        self.w_proj = capi.wproj_pars()
        self.a_proj = capi.AProjPars()
        self.pbeam_coefs = []
        self.s_kernel_function = "PSWF"
        self.s_fft_routine = "FFTW_ESTIMATE_FFT"
        self.s_interp_type = "linear"
        # source-find settings (load_json_config.h:209-226)
        self.detection_n_sigma = 0.0
        self.analysis_n_sigma = 0.0
        self.find_negative_sources = True
        self.estimate_rms = 0.0
        self.sigma_clip_iters = 5
        self.s_median_method = "BINMEDIAN"
        self.median_method = capi.MEDIAN_BINMEDIAN
        self.gaussian_fitting = True
        self.ccl_4connectivity = False
        self.generate_labelmap = False
        self.s_ceres_diffmethod = "AutoDiff_SingleResBlk"
        self.s_ceres_solvertype = "LinearSearch_BFGS"
        self.source_min_area = 5
        if document is None:
            with open(path) as f:
                document = json.load(f)
        if not isinstance(document, dict):
            raise ValueError("configuration must be a JSON object")
        self._imager(document.get("imager_settings"))
        self._wprojection(document.get("wprojection_settings"))
        self._aprojection(document.get("aprojection_settings"))
        self._sourcefind(document.get("sourcefind_settings"))

    def _imager(self, sec):
        p = self.img_pars
        if sec is None:
            return
        if "image_size_pix" in sec:
            p.image_size = int(sec["image_size_pix"])
        if "cell_size_arcsec" in sec:
            p.cell_size = float(sec["cell_size_arcsec"])
        if "padding_factor" in sec:
            p.padding_factor = float(sec["padding_factor"])
            p.padded_image_size = int(float(p.image_size) * p.padding_factor)
        else:
            p.padded_image_size = p.image_size
        if "kernel_function" in sec:
            self.s_kernel_function = sec["kernel_function"]
            p.kernel_function = _lookup(KERNEL_FUNCTIONS, self.s_kernel_function, "kernel function")
        for key, attr in (("kernel_support", "kernel_support"), ("oversampling", "oversampling")):
            if key in sec:
                setattr(p, attr, int(sec[key]))
        for key in ("kernel_exact", "generate_beam", "gridding_correction", "analytic_gcf"):
            if key in sec:
                setattr(p, key, int(bool(sec[key])))
        if "fft_routine" in sec:
            self.s_fft_routine = sec["fft_routine"]
            p.r_fft = _lookup(FFT_ROUTINES, self.s_fft_routine, "FFT routine")
        if "fft_wisdom_filename" in sec:
            self.fft_wisdom_filename = sec["fft_wisdom_filename"]

    def _wprojection(self, sec):
        if sec is None:
            return
        w = self.w_proj
        for key in ("num_wplanes", "max_wpconv_support", "undersampling_opt"):
            if key in sec:
                setattr(w, key, int(sec[key]))
        if "kernel_trunc_perc" in sec:
            w.kernel_trunc_perc = float(sec["kernel_trunc_perc"])
        for key in ("hankel_opt", "hankel_proj_slice", "wplanes_median"):
            if key in sec:
                setattr(w, key, int(bool(sec[key])))
        if "interp_type" in sec:
            self.s_interp_type = sec["interp_type"]
            w.interp_type = _lookup(INTERP_TYPES, self.s_interp_type, "interpolation type")

    def _aprojection(self, sec):
        if sec is None:
            return
        a = self.a_proj
        if "aproj_numtimesteps" in sec:
            a.num_timesteps = int(sec["aproj_numtimesteps"])
        for key in ("obs_dec", "obs_ra"):
            if key in sec:
                setattr(a, key, float(sec[key]))
        if "pbeam_coefs" in sec:
            self.pbeam_coefs = [float(x) for x in sec["pbeam_coefs"]]
        if "aproj_opt" in sec:
            a.aproj_opt = int(bool(sec["aproj_opt"]))
        if "aproj_mask_perc" in sec:
            a.aproj_mask_perc = float(sec["aproj_mask_perc"])

    def _sourcefind(self, sec):
        if sec is None:
            return
        if "sourcefind_detection" in sec:
            self.detection_n_sigma = float(sec["sourcefind_detection"])
        if "sourcefind_analysis" in sec:
            self.analysis_n_sigma = float(sec["sourcefind_analysis"])
        if "find_negative_sources" in sec:
            self.find_negative_sources = bool(sec["find_negative_sources"])
        if "rms_estimation" in sec:
            self.estimate_rms = float(sec["rms_estimation"])
        if "sigma_clip_iters" in sec:
            self.sigma_clip_iters = int(sec["sigma_clip_iters"])
        if "median_method" in sec:
            self.s_median_method = sec["median_method"]
            self.median_method = _lookup(MEDIAN_METHODS, self.s_median_method, "median method")
        for key in ("gaussian_fitting", "ccl_4connectivity", "generate_labelmap"):
            if key in sec:
                setattr(self, key, bool(sec[key]))
        if "source_min_area" in sec:
            self.source_min_area = int(sec["source_min_area"])
        if "ceres_diffmethod" in sec:
            self.s_ceres_diffmethod = _lookup(CERES_DIFF_METHODS, sec["ceres_diffmethod"], "Ceres diff method")
        if "ceres_solvertype" in sec:
            self.s_ceres_solvertype = _lookup(CERES_SOLVER_TYPES, sec["ceres_solvertype"], "Ceres solver type")

    def sourcefind_pars(self):
        """The SourceFindImage arguments of this configuration (reduce.cpp:130-132)."""
        return capi.sourcefind_pars(self.detection_n_sigma, self.analysis_n_sigma, rms_est=self.estimate_rms,
                                    find_negative_sources=self.find_negative_sources,
                                    sigma_clip_iters=self.sigma_clip_iters, median_method=self.median_method,
                                    gaussian_fitting=False, ccl_4connectivity=self.ccl_4connectivity,
                                    generate_labelmap=self.generate_labelmap, source_min_area=self.source_min_area)


def set_image_sizes(img_pars):
    """set_image_sizes (src/reduce/common_functions.cpp:49-78): image_size and the padded
    size rounded up to multiples of 4.  Returns the list of adjustments made."""
    notes = []
    n = img_pars.image_size
    while n % 4:
        n += 1
    if n != img_pars.image_size:
        notes.append(f"image size must be a multiple of four: set to {n} pixels")
        img_pars.image_size = n
    padded = int(float(img_pars.image_size) * img_pars.padding_factor)
    p = padded
    while p % 4:
        p += 1
    if p != padded:
        notes.append(f"padded image size must be a multiple of four: set to {p} pixels")
    img_pars.padded_image_size = p
    return notes


def _load(path, var_name):
    if var_name:
        with np.load(path, allow_pickle=False) as z:
            return np.array(z[var_name])
    return np.load(path, allow_pickle=False)


def _as_matrix(a):
    # column-major matrix of the array's shape; 1-D arrays become one column (load_data.h:33-49)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    elif a.ndim != 2:
        a = a.reshape(a.shape[0], -1)
    return np.asfortranarray(a)


def load_npy_complex_array(path, var_name=None):
    """load_npy_complex_array (load_data.h:19-78): complex64 or complex128 -> complex128."""
    a = _load(path, var_name)
    if a.dtype.kind != "c":
        raise ValueError("Invalid data type. Must be float or complex!")
    return _as_matrix(a.astype(np.complex128))


def load_npy_double_array(path, var_name=None):
    """load_npy_double_array (load_data.h:89-147): float32 or float64 -> float64."""
    a = _load(path, var_name)
    if a.dtype.kind != "f" or a.dtype.itemsize not in (4, 8):
        raise ValueError("Invalid data type. Must be float!")
    return _as_matrix(a.astype(np.float64))


def island_json(islands, fits=None):
    """The document save_json_sourcefind_output writes (src/auxiliary/save_json_sf_output.cpp:9-52):
    {"Islands": [{"sign", "val", "x_idx", "y_idx", "MomentsFit": {...}[, "GaussianFit": {...}]}]}."""
    names = ("amplitude", "x_centre", "y_centre", "semimajor", "semiminor", "theta")
    out = []
    for k, i in enumerate(islands):
        d = {"sign": int(i["sign"]), "val": float(i["extremum_val"]), "x_idx": int(i["extremum_x_idx"]),
             "y_idx": int(i["extremum_y_idx"]),
             "MomentsFit": {n: float(v) for n, v in zip(names, i["moments_fit"])}}
        if fits is not None:
            d["GaussianFit"] = {n: float(v) for n, v in zip(names, fits[k]["leastsq_fit"])}
        out.append(d)
    return {"Islands": out}
