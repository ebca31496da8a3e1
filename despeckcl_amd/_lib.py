"""Loader of libdespeckcl_amd.so (the HIP library) with ctypes signatures.

The library is built in-tree by despeckcl_amd.build; a missing library is an error, there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdespeckcl_amd.so")

DKX_OK = 0
DKX_EINVAL = -1
DKX_ERESOURCE = -2
DKX_ESTATS = -3
DKX_EUNSUPPORTED = -4
DKX_EHIP = -5


class DkxStats(C.Structure):
    _fields_ = [("patch_size", C.c_int), ("scale_size", C.c_int), ("lut_size", C.c_int),
                ("dissims_min", C.c_float), ("dissims_max", C.c_float),
                ("quantilles", C.POINTER(C.c_float)), ("chi2cdf_inv", C.POINTER(C.c_float))]


class DkxImage(C.Structure):
    _fields_ = [("ampl_master", C.c_void_p), ("ampl_slave", C.c_void_p), ("phase", C.c_void_p),
                ("ref_filt", C.c_void_p), ("phase_filt", C.c_void_p), ("coh_filt", C.c_void_p),
                ("height", C.c_int), ("width", C.c_int)]


class DkxCovmatImage(C.Structure):
    _fields_ = [("covmat", C.c_void_p), ("covmat_filt", C.c_void_p), ("height", C.c_int), ("width", C.c_int)]


class DkxRequest(C.Structure):
    _fields_ = [("kind", C.c_int), ("ampl_master", C.c_void_p), ("ampl_slave", C.c_void_p),
                ("phase", C.c_void_p), ("ref_filt", C.c_void_p), ("phase_filt", C.c_void_p),
                ("coh_filt", C.c_void_p), ("height", C.c_int), ("width", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_IP = C.POINTER(C.c_int)
_SP = C.POINTER(DkxStats)
_ULLP = C.POINTER(C.c_ulonglong)

# name -> (restype, argtypes); every symbol declared in include/despeckcl_amd.h
SIGNATURES = {
    "dkx_last_error": (C.c_char_p, []),
    "dkx_version": (_I, []),
    "dkx_nlsar_insar": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F]),
    "dkx_nlsar_insar_device": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _IP, _I, _IP, _I, _SP, _I,
                                    _F, _F, _P, _P, _ULLP, _P]),
    "dkx_nlsar_insar_device_noalloc": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _IP, _I, _IP, _I, _SP,
                                            _I, _F, _F, _P, _P, _ULLP, _P]),
    "dkx_nlsar_covmat": (_I, [_P, _P, _I, _I, _I, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F]),
    "dkx_nlsar_covmat_device": (_I, [_P, _P, _I, _I, _I, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F,
                                     _P, _P, _ULLP, _P]),
    "dkx_nlinsar": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I]),
    "dkx_set_nlinsar_weight_store": (_I, [_I]),
    "dkx_get_nlinsar_weight_store": (_I, []),
    "dkx_boxcar": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I]),
    "dkx_goldstein": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F]),
    "dkx_goldstein_device": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _P]),
    "dkx_boxcar_device": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _P]),
    "dkx_nlinsar_device": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _ULLP, _P]),
    "dkx_nlsar_training_insar": (_I, [_P, _P, _P, _I, _I, _IP, _I, _IP, _I, _SP, _P]),
    "dkx_nlsar_training_covmat": (_I, [_P, _I, _I, _I, _IP, _I, _IP, _I, _SP, _P]),
    "dkx_nlsar_training_covmat_device": (_I, [_P, _I, _I, _I, _IP, _I, _IP, _I, _SP, _P, _P]),
    "dkx_nlsar_stats_build": (C.c_longlong, [_P, C.c_longlong, _I, _P, _P, _P, _P]),
    "dkx_chi2_inv": (C.c_double, [C.c_double, C.c_double]),
    "dkx_nlsar_stats_store": (_I, [C.c_char_p, _SP, _I]),
    "dkx_nlsar_stats_load": (C.c_longlong, [C.c_char_p, _SP, _I, _P, C.c_longlong,
                                            C.POINTER(C.c_longlong)]),
    "dkx_set_tile": (_I, [_I, _I]),
    "dkx_get_tile": (_I, [_I]),
    "dkx_set_timing": (_I, [_I]),
    "dkx_stage_times": (_I, [_P, _P, _I]),
    "dkx_release_workspace": (None, []),
    "dkx_nlsar_batch_create": (_I, [C.POINTER(C.c_void_p)]),
    "dkx_nlsar_batch_destroy": (None, [_P]),
    "dkx_nlsar_batch_set": (_I, [_P, C.POINTER(DkxImage), _I, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F]),
    "dkx_nlsar_batch_set_covmat": (_I, [_P, C.POINTER(DkxCovmatImage), _I, _I, _I, _IP, _I, _IP, _I, _SP, _I, _F,
                                        _F]),
    "dkx_nlsar_batch_enqueue": (_I, [_P, _P]),
    "dkx_nlsar_batch_graph_launch": (_I, [_P, _P]),
    "dkx_nlsar_batch_set_tile": (_I, [_P, _I]),
    "dkx_nlsar_batch_set_diagnostics": (_I, [_P, _I]),
    "dkx_nlsar_batch_get_tile": (_I, [_P]),
    "dkx_nlsar_batch_autotune": (_I, [_P, _P, _I, _IP, _P]),
    "dkx_nlsar_batch_ub_pixels": (_I, [_P, _ULLP]),
    "dkx_nlsar_insar_device_autotune": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F,
                                             _I, _I, _P, _IP, _P]),
    "dkx_nlinsar_device_autotune": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _IP, _P]),
    "dkx_get_tuned_tile": (_I, [_I, _I, _I, _I, _IP, _I, _IP, _I, _I]),
    "dkx_clear_tuned_tiles": (None, []),
    "dkx_nlinsar_simulate_h": (_I, [_I, C.c_uint, C.POINTER(C.c_float)]),
    "dkx_nlinsar_batch_create": (_I, [C.POINTER(C.c_void_p)]),
    "dkx_nlinsar_batch_destroy": (None, [_P]),
    "dkx_nlinsar_batch_set": (_I, [_P, C.POINTER(DkxImage), _I, _I, _I, _I, _I]),
    "dkx_nlinsar_batch_enqueue": (_I, [_P, _P]),
    "dkx_nlinsar_batch_graph_launch": (_I, [_P, _P]),
    "dkx_nlinsar_batch_smoothed": (_I, [_P, _ULLP]),
    "dkx_server_create": (_I, [C.c_ulonglong, _I, _IP, _I, _IP, _I, _SP, _I, _F, _F, _I, _I, _I, _I,
                               C.POINTER(C.c_void_p)]),
    "dkx_server_submit": (_I, [_P, C.POINTER(DkxRequest), C.c_double, C.POINTER(C.c_longlong)]),
    "dkx_server_poll": (_I, [_P, C.c_longlong, C.POINTER(C.c_double)]),
    "dkx_server_wait": (_I, [_P, C.c_longlong, C.POINTER(C.c_double)]),
    "dkx_server_run_trace": (_I, [_P, C.POINTER(DkxRequest), C.POINTER(C.c_double), _I,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "dkx_server_counters": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "dkx_server_nlinsar_counters": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "dkx_server_destroy": (None, [_P]),
}

_lib = None


class DespeckleError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m despeckcl_amd.build` "
                "(the filters run only on the HIP library, there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc == DKX_OK:
        return
    msg = lib().dkx_last_error().decode(errors="replace")
    if rc in (DKX_EINVAL, DKX_ESTATS, DKX_EUNSUPPORTED):
        raise ValueError(msg)
    if rc == DKX_ERESOURCE:
        raise MemoryError(msg)
    raise DespeckleError(msg)


def int_array(values):
    values = [int(v) for v in values]
    return (C.c_int * max(1, len(values)))(*values), len(values)
