"""ctypes binding of libnngp_hip.so (the C ABI of include/nngp.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

# NNGP_LIB names another build of the same library (the host-sanitized
# copy of `python -m spnngp_amd.build --sanitize`); by default the in-tree
# libnngp_hip.so
LIB_PATH = os.environ.get("NNGP_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libnngp_hip.so")

NNGP_OK = 0
NNGP_ERR_ARG = 1
NNGP_ERR_NOT_PD = 2
NNGP_ERR_COV_MODEL = 3
NNGP_ERR_HIP = 4
NNGP_ERR_NOMEM = 5
NNGP_ERR_INTERRUPT = 6
NNGP_ERR_UNSUPPORTED = 7
NNGP_ERR_TIMEOUT = 8

NNGP_FP64 = 0
NNGP_MIXED = 1
TUNE_MAX = 12  # NNGP_TUNE_MAX of include/nngp.h

# Every symbol declared in include/nngp.h.
EXPORTS = (
    "nngp_last_error", "nngp_version", "nngp_device_count",
    "nngp_mkNNIndxCB", "nngp_mkNNIndx", "nngp_nn2",
    "nngp_updateBF", "nngp_updateConjBF", "nngp_Q",
    "nngp_rNNGP", "nngp_cNNGP", "nngp_rNNGPPredict",
    "nngp_model_create", "nngp_model_destroy", "nngp_model_update",
    "nngp_model_qresid", "nngp_model_get_bf", "nngp_model_rNNGP",
    "nngp_prior_draw", "nngp_model_profile", "nngp_model_timing",
    "nngp_model_loop_timing",
    "nngp_server_create", "nngp_server_destroy", "nngp_server_submit",
    "nngp_server_wait", "nngp_server_release", "nngp_server_predict",
    "nngp_server_get_stats",
    "nngp_cvgrid_create", "nngp_cvgrid_run", "nngp_cvgrid_destroy",
    "nngp_model_autotune", "nngp_model_set_shape", "nngp_rNNGPReplicated",
    "nngp_rng_r_create", "nngp_rng_r_destroy", "nngp_mkUIndx", "nngp_sNNGP",
    "nngp_sNNGPPredict", "nngp_Q_noalloc",
)

_RNG_FN = C.CFUNCTYPE(C.c_double, C.c_void_p)
_INT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class NngpRng(C.Structure):
    _fields_ = [("state", C.c_void_p), ("norm_rand", _RNG_FN),
                ("unif_rand", _RNG_FN)]


class NngpOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int),
                ("seed", C.c_ulonglong), ("interrupt", _INT_FN),
                ("interrupt_arg", C.c_void_p), ("shape", C.c_int * 3),
                ("sweep", C.c_int), ("search", C.c_int),
                ("reserved", C.c_int * 3)]


class NngpPassOut(C.Structure):
    _fields_ = [("logDet", C.c_double), ("qResid", C.c_double),
                ("XtViX", C.c_double * 64), ("XtViy", C.c_double * 8),
                ("ytViy", C.c_double)]


class NngpServerOpts(C.Structure):
    _fields_ = [("device_bytes", C.c_ulonglong), ("host_bytes", C.c_ulonglong),
                ("batch_sites", C.c_int), ("depth", C.c_int),
                ("device", C.c_int), ("reserved", C.c_int * 8)]


class NngpRequestInfo(C.Structure):
    _fields_ = [("q", C.c_int), ("nSamples", C.c_int), ("status", C.c_int),
                ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("y0", C.POINTER(C.c_double))]


class NngpServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("launches", C.c_longlong), ("pending", C.c_int),
                ("active", C.c_int), ("site_samples", C.c_double),
                ("device_bytes_used", C.c_ulonglong),
                ("device_bytes_peak", C.c_ulonglong),
                ("host_bytes_used", C.c_ulonglong),
                ("host_bytes_peak", C.c_ulonglong),
                ("kernel_ms", C.c_double),
                ("kernel_site_samples", C.c_double),
                ("done_unreported", C.c_longlong)]


class NngpCvGridDetail(C.Structure):
    _fields_ = [("betaHat", C.c_void_p), ("ab", C.c_void_p),
                ("bBInv", C.c_void_p), ("y0Hat", C.c_void_p),
                ("y0HatVar", C.c_void_p), ("graph_ms", C.c_double),
                ("bf_family", C.c_int), ("bf_threads", C.c_int),
                ("bf_mode", C.c_int),
                ("tune_n", C.c_int), ("tune_family", C.c_int * TUNE_MAX),
                ("tune_threads", C.c_int * TUNE_MAX),
                ("tune_mode", C.c_int * TUNE_MAX),
                ("tune_ms", C.c_double * TUNE_MAX)]


class NngpTuneReport(C.Structure):
    _fields_ = [("n", C.c_int), ("family", C.c_int * TUNE_MAX),
                ("threads", C.c_int * TUNE_MAX), ("mode", C.c_int * TUNE_MAX),
                ("ms", C.c_double * TUNE_MAX), ("best_family", C.c_int),
                ("best_threads", C.c_int), ("best_mode", C.c_int)]


class NngpError(RuntimeError):
    """A nonzero status from libnngp_hip.so."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i = C.c_int
_d = C.c_double

_lib = None


def load():
    """Load libnngp_hip.so once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m "
            "spnngp_amd.build` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    L.nngp_last_error.restype = C.c_char_p
    L.nngp_version.restype = _i
    L.nngp_device_count.restype = _i
    for f in (L.nngp_mkNNIndxCB, L.nngp_mkNNIndx):
        f.argtypes = [_i, _i, _dp, _ip, _dp, _ip, _i]
    L.nngp_nn2.argtypes = [_i, _dp, _i, _dp, _i, _ip, _vp]
    L.nngp_updateBF.argtypes = [_dp, _dp, _dp, _ip, _ip, _i, _i, _dp, _i,
                                C.POINTER(_d), C.POINTER(NngpOpts)]
    L.nngp_updateConjBF.argtypes = [_dp, _dp, _dp, _ip, _ip, _i, _i, _d, _d,
                                    _d, _i, C.POINTER(NngpOpts)]
    L.nngp_Q_noalloc.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp,
                                 _vp]
    L.nngp_Q.argtypes = [_dp, _dp, _dp, _dp, _i, _i, _ip, _ip,
                         C.POINTER(_d)]
    L.nngp_rNNGP.argtypes = (
        [_dp, _dp, _i, _i, _i, _dp, _i, _ip, _ip, _dp, _dp, _dp, _vp, _dp]
        + [_d] * 8
        + [_i, _i, _i, _i, _i, _vp, C.POINTER(NngpRng), C.POINTER(NngpOpts),
           _dp, _dp, _vp, C.POINTER(_i)])
    L.nngp_cNNGP.argtypes = [_dp, _dp, _i, _i, _dp, _dp, _ip, _ip, _vp, _vp,
                             _i, _vp, _i, _i, _dp, _i, _i, _i,
                             C.POINTER(NngpOpts), _dp, _dp, _dp, _dp, _vp,
                             _vp]
    L.nngp_rNNGPPredict.argtypes = [_dp, _dp, _dp, _i, _i, _i, _dp, _dp, _i,
                                    _ip, _dp, _dp, _i, _i, _i, _i, _i, _vp,
                                    C.POINTER(NngpRng), C.POINTER(NngpOpts),
                                    _dp]
    L.nngp_model_create.argtypes = [C.POINTER(_vp), _i, _i, _i, _dp, _ip,
                                    _ip, _vp, _vp, _i, C.POINTER(NngpOpts)]
    L.nngp_model_destroy.argtypes = [_vp]
    L.nngp_model_destroy.restype = None
    L.nngp_model_update.argtypes = [_vp, _dp, _i, _vp, _i,
                                    C.POINTER(NngpPassOut)]
    L.nngp_model_qresid.argtypes = [_vp, _i, _dp, C.POINTER(_d)]
    L.nngp_model_get_bf.argtypes = [_vp, _i, _vp, _vp]
    L.nngp_model_rNNGP.argtypes = (
        [_vp, _dp, _dp, _dp, _vp, _dp] + [_d] * 8
        + [_i, _i, _i, C.POINTER(NngpRng), C.POINTER(NngpOpts), _dp, _dp,
           C.POINTER(_i)])
    L.nngp_prior_draw.argtypes = [_dp, _dp, _ip, _ip, _i, _dp, _dp]
    L.nngp_rNNGPReplicated.argtypes = [_vp, _i, _i, _i, _dp, _i, _ip, _ip,
                                       _vp, _dp, _i, _i, _i, _vp, _vp, _dp]
    L.nngp_rng_r_create.argtypes = [C.c_uint, C.POINTER(NngpRng)]
    L.nngp_rng_r_destroy.argtypes = [C.POINTER(NngpRng)]
    L.nngp_rng_r_destroy.restype = None
    L.nngp_sNNGPPredict.argtypes = [_dp, _vp, _dp, _i, _i, _i, _dp, _dp, _i,
                                    _ip, _dp, _dp, _dp, _i, _i, _i, _i, _i,
                                    _i, _vp, C.POINTER(NngpRng),
                                    C.POINTER(NngpOpts), _dp, _dp]
    L.nngp_mkUIndx.argtypes = [_i, _i, _ip, _ip, _ip, _ip, _ip, _i]
    L.nngp_sNNGP.argtypes = (
        [_dp, _dp, _i, _i, _i, _dp, _i, _ip, _ip, _ip, _ip, _ip, _dp, _dp,
         _dp, _vp, _dp] + [_d] * 8
        + [_i, _i, _i, _i, C.POINTER(NngpRng), C.POINTER(NngpOpts), _dp, _dp,
           _dp, C.POINTER(_i)])
    L.nngp_model_profile.argtypes = [_vp, _i]
    L.nngp_model_timing.argtypes = [_vp, _dp]
    L.nngp_model_loop_timing.argtypes = [_vp, _dp]
    L.nngp_server_create.argtypes = [C.POINTER(_vp), _dp, _dp, _dp, _i, _i,
                                     _i, _dp, _dp, _i, _i,
                                     C.POINTER(NngpServerOpts)]
    L.nngp_server_destroy.argtypes = [_vp]
    L.nngp_server_destroy.restype = None
    L.nngp_server_submit.argtypes = [_vp, _i, _dp, _dp, C.c_ulonglong,
                                     C.POINTER(C.c_longlong)]
    L.nngp_server_wait.argtypes = [_vp, C.POINTER(C.c_longlong), _d,
                                   C.POINTER(NngpRequestInfo)]
    L.nngp_server_release.argtypes = [_vp, C.c_longlong]
    L.nngp_server_predict.argtypes = [_vp, _i, _dp, _dp, C.c_ulonglong, _dp]
    L.nngp_server_get_stats.argtypes = [_vp, C.POINTER(NngpServerStats)]
    L.nngp_cvgrid_create.argtypes = [C.POINTER(_vp), _dp, _dp, _i, _i, _dp,
                                     _i, _i, _ip, _i, _dp, _i,
                                     C.POINTER(NngpOpts)]
    L.nngp_cvgrid_run.argtypes = [_vp, _dp, _i, _dp,
                                  C.POINTER(NngpCvGridDetail)]
    L.nngp_cvgrid_destroy.argtypes = [_vp]
    L.nngp_model_autotune.argtypes = [_vp, _dp, C.POINTER(NngpTuneReport)]
    L.nngp_model_set_shape.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.nngp_cvgrid_destroy.restype = None
    _lib = L
    return _lib


def check(status: int):
    if status != NNGP_OK:
        msg = load().nngp_last_error()
        raise NngpError(status, msg.decode() if msg else "")


def exported_symbols():
    """Names of include/nngp.h functions the loaded library exports."""
    L = load()
    found = []
    for name in EXPORTS:
        try:
            getattr(L, name)
            found.append(name)
        except AttributeError:
            pass
    return found
