"""ctypes binding of libgwl.so (the C ABI declared in include/gwl.h).

The library is built in-tree by ``python -m bilby_amd.build``.  There is no
CPU fallback: if the library is missing, importing this module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgwl.so")

from .abi import *  # noqa: F401,F403
from .abi import NPARAM, PARAM_COLUMNS  # noqa: F401

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class EventDesc(ctypes.Structure):
    _fields_ = [
        ("n_det", ctypes.c_int32), ("n_freq", ctypes.c_int32),
        ("duration", ctypes.c_double), ("sampling_frequency", ctypes.c_double),
        ("start_time", ctypes.c_double),
        ("mask_lo", _ip), ("mask_hi", _ip),
        ("strain", _dp), ("psd", _dp), ("detector_tensor", _dp), ("vertex", _dp),
        ("approximant", ctypes.c_int32),
        ("wf_minimum_frequency", ctypes.c_double), ("wf_maximum_frequency", ctypes.c_double),
        ("reference_frequency", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("time_prior_minimum", ctypes.c_double), ("time_prior_maximum", ctypes.c_double),
        ("reference_distance", ctypes.c_double),
    ]


class RoqDesc(ctypes.Structure):
    _fields_ = [
        ("n_det", ctypes.c_int32), ("n_linear", ctypes.c_int32),
        ("n_quadratic", ctypes.c_int32), ("n_time_samples", ctypes.c_int32),
        ("start_time", ctypes.c_double), ("time_samples_start", ctypes.c_double),
        ("time_samples_step", ctypes.c_double),
        ("linear_nodes", _dp), ("quadratic_nodes", _dp), ("linear_weights", _dp),
        ("quadratic_weights", _dp), ("detector_tensor", _dp), ("vertex", _dp),
        ("reference_frequency", ctypes.c_double), ("wf_minimum_frequency", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("time_prior_minimum", ctypes.c_double), ("time_prior_maximum", ctypes.c_double),
        ("delta_tc", ctypes.c_double), ("reference_distance", ctypes.c_double),
        ("time_samples", _dp),
    ]


class RelbinDesc(ctypes.Structure):
    _fields_ = [
        ("n_bins", ctypes.c_int32), ("bin_frequencies", _dp), ("fiducial_response", _dp),
        ("summary", _dp),
    ]


class MbTimeDesc(ctypes.Structure):
    """gwl_mb_time_desc (include/gwl.h)."""
    _fields_ = [("grid_index", ctypes.POINTER(ctypes.c_int64)), ("n_fft", ctypes.c_int64),
                ("j_first", ctypes.c_int64), ("n_window", ctypes.c_int32),
                ("t_step", ctypes.c_double), ("antenna_time", ctypes.c_double)]


class MbFftBand(ctypes.Structure):
    """gwl_mb_fft_band (include/gwl.h)."""
    _fields_ = [("half_length", ctypes.c_int32), ("first_point", ctypes.c_int32),
                ("n_points", ctypes.c_int32), ("first_bin", ctypes.c_int32),
                ("sqrt_window", ctypes.POINTER(ctypes.c_double)),
                ("odd_weights", ctypes.POINTER(ctypes.c_double))]


class GWLError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libgwl.so is not built (%s). Run `python -m bilby_amd.build` first; "
            "there is no CPU fallback for the likelihood." % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "gwl_last_error": ([], ctypes.c_char_p),
        "gwl_ctx_create": ([ctypes.c_int32, ctypes.POINTER(vp)], ctypes.c_int),
        "gwl_ctx_destroy": ([vp], ctypes.c_int),
        "gwl_event_create": ([vp, ctypes.POINTER(EventDesc), ctypes.POINTER(vp),
                              ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_event_destroy": ([vp], ctypes.c_int),
        "gwl_event_set_distance_spline": ([vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _dp],
                                          ctypes.c_int),
        "gwl_event_reserve": ([vp, ctypes.c_int64], ctypes.c_int),
        "gwl_eval": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp, vp],
                     ctypes.c_int),
        "gwl_eval_noalloc": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp, vp],
                             ctypes.c_int),
        "gwl_detector_response": ([vp, _dp, ctypes.c_int32, ctypes.c_int32, _dp], ctypes.c_int),
        "gwl_distance_table": ([vp, _dp, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                _dp, ctypes.c_int32, _dp, ctypes.c_int32, _dp], ctypes.c_int),
        "gwl_spline2d_coefficients": ([ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, _dp],
                                      ctypes.c_int),
        "gwl_spline2d_eval_host": ([ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, ctypes.c_int64,
                                    _dp, _dp, _dp], ctypes.c_int),
        "gwl_fft_batch": ([vp, vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp],
                          ctypes.c_int),
        "gwl_ln_i0": ([vp, vp, vp, ctypes.c_int64, vp], ctypes.c_int),
        "gwl_noise_log_likelihood": ([vp, _dp], ctypes.c_int),
        "gwl_event_device_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_event_timing": ([vp, ctypes.c_int32], ctypes.c_int),
        "gwl_event_set_tile": ([vp, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "gwl_roq_create": ([vp, ctypes.POINTER(RoqDesc), ctypes.POINTER(vp),
                            ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_roq_destroy": ([vp], ctypes.c_int),
        "gwl_roq_set_distance_spline": ([vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _dp],
                                        ctypes.c_int),
        "gwl_roq_eval": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp],
                         ctypes.c_int),
        "gwl_multi_create": ([vp, ctypes.c_int32, ctypes.POINTER(vp), ctypes.POINTER(vp),
                              ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(vp)],
                             ctypes.c_int),
        "gwl_multi_set_rows": ([vp, ctypes.POINTER(ctypes.c_int64), vp], ctypes.c_int),
        "gwl_multi_set_tile": ([vp, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "gwl_multi_launch": ([vp, vp], ctypes.c_int),
        "gwl_multi_capture": ([vp, vp], ctypes.c_int),
        "gwl_multi_destroy": ([vp], ctypes.c_int),
        "gwl_gemm": ([vp, ctypes.c_int32, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32,
                      ctypes.c_int32, vp], ctypes.c_int),
        "gwl_event_timing_read": ([vp, _dp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_roq_reserve": ([vp, ctypes.c_int64], ctypes.c_int),
        "gwl_event_set_transform": ([vp, ctypes.c_int32], ctypes.c_int),
        "gwl_event_set_escalation": ([vp, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "gwl_event_set_multi_steps": ([vp, ctypes.c_int32], ctypes.c_int),
        "gwl_event_set_polarizations": ([vp, vp, ctypes.c_int64], ctypes.c_int),
        "gwl_event_set_time_prior": ([vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp],
                                     ctypes.c_int),
        "gwl_roq_set_time_prior": ([vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp],
                                   ctypes.c_int),
        "gwl_event_escalated_rows": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_server_create": ([vp, ctypes.c_int64, ctypes.POINTER(vp)], ctypes.c_int),
        "gwl_server_destroy": ([vp], ctypes.c_int),
        "gwl_server_add_event": ([vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
                                 ctypes.c_int),
        "gwl_server_add_roq": ([vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
                               ctypes.c_int),
        "gwl_server_add_group": ([vp, ctypes.c_int32, ctypes.POINTER(vp),
                                  ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "gwl_server_submit": ([vp, ctypes.c_int32, _dp, ctypes.c_int64, _dp,
                               ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_server_wait": ([vp, ctypes.c_int64, _dp], ctypes.c_int),
        "gwl_server_stats": ([vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
                             ctypes.c_int),
        "gwl_server_device_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_server_listen": ([vp, ctypes.c_char_p], ctypes.c_int),
        "gwl_relbin_create": ([vp, ctypes.POINTER(RelbinDesc), ctypes.POINTER(vp)], ctypes.c_int),
        "gwl_relbin_destroy": ([vp], ctypes.c_int),
        "gwl_relbin_reserve": ([vp, ctypes.c_int64], ctypes.c_int),
        "gwl_roq_weights_linear": ([vp, ctypes.c_int32, ctypes.c_int64, _dp,
                                    ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int32, ctypes.c_double, _dp], ctypes.c_int),
        "gwl_mb_create": ([vp, ctypes.c_int32, _dp, _dp, _dp, ctypes.POINTER(vp)], ctypes.c_int),
        "gwl_mb_destroy": ([vp], ctypes.c_int),
        "gwl_mb_set_time": ([vp, ctypes.POINTER(MbTimeDesc)], ctypes.c_int),
        "gwl_mb_set_fft": ([vp, ctypes.c_int32, ctypes.POINTER(MbFftBand)], ctypes.c_int),
        "gwl_eval_time_series": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp,
                                  vp], ctypes.c_int),
        "gwl_event_time_window": ([vp, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "gwl_relbin_set_time": ([vp, ctypes.POINTER(ctypes.c_int64), _dp], ctypes.c_int),
        "gwl_relbin_time_series": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp,
                                    vp, vp], ctypes.c_int),
        "gwl_relbin_set_upsampled": ([vp, ctypes.POINTER(ctypes.c_int64), _dp, _dp,
                                      ctypes.c_int64], ctypes.c_int),
        "gwl_relbin_upsampled_series": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                         ctypes.c_int32, vp, vp, vp], ctypes.c_int),
        "gwl_weighted_inner_products": ([vp, _dp, _dp, _dp], ctypes.c_int),
        "gwl_window_time_series": ([vp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _dp,
                                    _dp], ctypes.c_int),
        "gwl_mb_time_series": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp,
                                vp], ctypes.c_int),
        "gwl_roq_inner_products": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, vp, vp, vp], ctypes.c_int),
        "gwl_mb_reserve": ([vp, ctypes.c_int64], ctypes.c_int),
        "gwl_mb_device_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_mb_eval": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp, vp],
                        ctypes.c_int),
        "gwl_calib_create": ([vp, ctypes.c_int32, _dp, ctypes.POINTER(vp)], ctypes.c_int),
        "gwl_calib_destroy": ([vp], ctypes.c_int),
        "gwl_calib_reserve": ([vp, ctypes.c_int64], ctypes.c_int),
        "gwl_calib_device_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_calib_eval": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp, vp],
                           ctypes.c_int),
        "gwl_server_add_calib": ([vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
                                 ctypes.c_int),
        "gwl_server_add_relbin": ([vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
                                  ctypes.c_int),
        "gwl_relbin_device_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "gwl_relbin_eval": ([vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp, vp],
                            ctypes.c_int),
    }
    for name, (argt, rest) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    return lib


LIB = _load()


def check(rc, what="libgwl"):
    if rc != 0:
        msg = LIB.gwl_last_error().decode(errors="replace")
        raise GWLError("%s failed (%d): %s" % (what, rc, msg))


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


EXPORTED_SYMBOLS = [
    "gwl_ctx_create", "gwl_ctx_destroy", "gwl_last_error", "gwl_event_create",
    "gwl_event_destroy", "gwl_event_set_distance_spline", "gwl_event_reserve", "gwl_eval",
    "gwl_detector_response", "gwl_distance_table", "gwl_spline2d_coefficients",
    "gwl_spline2d_eval_host", "gwl_fft_batch", "gwl_ln_i0", "gwl_noise_log_likelihood",
    "gwl_event_timing", "gwl_event_timing_read", "gwl_event_set_tile",
    "gwl_roq_create", "gwl_roq_destroy", "gwl_roq_set_distance_spline", "gwl_roq_eval",
    "gwl_gemm", "gwl_multi_create", "gwl_multi_set_rows", "gwl_multi_set_tile",
    "gwl_multi_launch", "gwl_multi_capture", "gwl_multi_destroy", "gwl_event_device_bytes",
    "gwl_roq_reserve", "gwl_server_create", "gwl_server_destroy", "gwl_server_add_event",
    "gwl_server_add_roq", "gwl_server_add_group", "gwl_server_submit", "gwl_server_wait",
    "gwl_server_stats", "gwl_server_device_bytes", "gwl_event_set_transform",
    "gwl_server_listen", "gwl_relbin_create", "gwl_relbin_destroy", "gwl_relbin_device_bytes",
    "gwl_relbin_eval", "gwl_server_add_relbin", "gwl_server_add_calib", "gwl_relbin_reserve", "gwl_calib_create",
    "gwl_calib_destroy", "gwl_calib_reserve", "gwl_calib_device_bytes", "gwl_calib_eval",
    "gwl_mb_create", "gwl_mb_destroy", "gwl_mb_reserve", "gwl_mb_device_bytes", "gwl_mb_eval",
    "gwl_mb_set_time", "gwl_mb_set_fft", "gwl_relbin_set_time", "gwl_eval_time_series", "gwl_event_time_window",
    "gwl_roq_weights_linear", "gwl_relbin_time_series", "gwl_mb_time_series",
    "gwl_roq_inner_products", "gwl_relbin_set_upsampled", "gwl_relbin_upsampled_series",
    "gwl_weighted_inner_products", "gwl_window_time_series", "gwl_event_set_escalation",
    "gwl_event_escalated_rows", "gwl_eval_noalloc", "gwl_event_set_time_prior",
    "gwl_roq_set_time_prior", "gwl_event_set_multi_steps", "gwl_event_set_polarizations",
]
