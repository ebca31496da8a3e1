"""Device-buffer entry points (inputs already in HBM), for callers that manage their own
device memory and streams, e.g. with torch tensors on a ROCm device.

Pointers are integers (tensor.data_ptr()); `stream` is a hipStream_t as an integer
(torch.cuda.current_stream().cuda_stream) or 0 for the default stream.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, int_array, lib


def nlsar_insar(a1, a2, phase, ref, phase_filt, coh, height, width, search_window_size,
                patch_sizes, scale_sizes, nlsar_stats, h=15.0, c=49.0, enl_maps=0, alpha_maps=0,
                stream=0, want_ub=False):
    """Enqueue NL-SAR on device buffers.  Returns the number of pixels with an undefined
    best-parameter selection when want_ub (synchronises), else None."""
    from . import _stats_structs
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    arr, ns, keep = _stats_structs(nlsar_stats)
    ub = C.c_ulonglong(0)
    check(lib().dkx_nlsar_insar_device(a1, a2, phase, ref, phase_filt, coh, int(height), int(width),
                                       int(search_window_size), P, nP, S, nS, arr, ns, float(h),
                                       float(c), enl_maps or None, alpha_maps or None,
                                       C.byref(ub) if want_ub else None, stream or None))
    return ub.value if want_ub else None


def nlinsar(a1, a2, phase, ref, phase_filt, coh, height, width, search_window_size, patch_size,
            niter, lmin, stream=0, want_count=False):
    """Enqueue NL-InSAR on device buffers (ref/phase_filt/coh hold the initial estimates).
    Returns the number of smoothed pixel-iterations when want_count (synchronises)."""
    n = C.c_ulonglong(0)
    check(lib().dkx_nlinsar_device(a1, a2, phase, ref, phase_filt, coh, int(height), int(width),
                                   int(search_window_size), int(patch_size), int(niter), int(lmin),
                                   C.byref(n) if want_count else None, stream or None))
    return n.value if want_count else None


def set_tile(filter_index, tile_index):
    check(lib().dkx_set_tile(int(filter_index), int(tile_index)))


def get_tile(filter_index):
    return lib().dkx_get_tile(int(filter_index))
