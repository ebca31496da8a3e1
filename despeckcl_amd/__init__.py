"""MI355X NL-SAR / NL-InSAR despeckling with the Python API of despeckCL's `despeckcl` module.

The functions below keep the names, argument order, defaults and return values of the
reference's SWIG module (swig/python/nlsar.i:45-258, swig/python/nlinsar.i:36-74), so
`import despeckcl_amd as despeckcl` is a drop-in for the nonlocal filters.  All filtering
runs in libdespeckcl_amd.so (hand-written HIP kernels for gfx950) through its C ABI
(include/despeckcl_amd.h).
"""
from __future__ import annotations

import ctypes as C
import os
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import DespeckleError, DkxStats, check, int_array, lib

__all__ = ["params", "stats", "nlsar_stats_collection", "nlsar_stats_collections",
           "nlsar_insar", "nlsar", "nlsar_train", "nlsar_train_insar", "nlinsar", "boxcar", "goldstein",
           "store_nlsar_stats_collection", "load_nlsar_stats_collection", "DespeckleError"]

_DEFAULT_LOG_LEVELS = ["error", "warning", "fatal"]  # nlsar.i:164, nlinsar.i:43

# key of the statistics map: ordered by patch size, then scale size (parameters.h:37-40)
params = namedtuple("params", ["patch_size", "scale_size"])


class stats:
    """Dissimilarity statistics of one (patch, scale) pair (include/compute/nlsar/stats.h:27-55)."""

    __slots__ = ("lut_size", "dissims_min", "dissims_max", "quantilles", "chi2cdf_inv")

    def __init__(self, dissims=None, lut_size=256):
        self.lut_size = int(lut_size)
        self.dissims_min = 0.0
        self.dissims_max = 0.0
        self.quantilles = np.zeros(self.lut_size, np.float32)
        self.chi2cdf_inv = np.zeros(self.lut_size, np.float32)
        if dissims is not None:
            d = np.ascontiguousarray(dissims, dtype=np.float32)
            dmin, dmax = C.c_float(), C.c_float()
            n = lib().dkx_nlsar_stats_build(d.ctypes.data, d.size, self.lut_size, C.byref(dmin),
                                            C.byref(dmax), self.quantilles.ctypes.data,
                                            self.chi2cdf_inv.ctypes.data)
            if n < 0:
                check(int(n))
            if n == 0:
                raise ValueError("stats: no finite dissimilarity")
            self.dissims_min, self.dissims_max = dmin.value, dmax.value

    def __repr__(self):
        return (f"stats(lut_size={self.lut_size}, dissims_min={self.dissims_min}, "
                f"dissims_max={self.dissims_max})")


class nlsar_stats_collection(dict):
    """std::map<nlsar::params, nlsar::stats>: a dict keyed by params(patch_size, scale_size),
    iterated in the map's key order."""

    def __iter__(self):
        return iter(sorted(dict.keys(self)))

    def keys(self):
        return sorted(dict.keys(self))

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def values(self):
        return [self[k] for k in self.keys()]


nlsar_stats_collections = nlsar_stats_collection  # SWIG template name, nlsar.i:13


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _same_shape(*arrays):
    shp = arrays[0].shape
    if len(shp) != 2 or any(a.shape != shp for a in arrays):
        raise ValueError("inputs must be 2-D arrays of one shape")
    return shp


def _stats_structs(nlsar_stats):
    """ctypes array of DkxStats for a collection; the returned list keeps the buffers alive."""
    items = sorted(nlsar_stats.items(), key=lambda kv: (kv[0][0], kv[0][1]))
    arr = (DkxStats * max(1, len(items)))()
    keep = []
    for i, (k, s) in enumerate(items):
        q = _f32(s.quantilles)
        c = _f32(s.chi2cdf_inv)
        keep += [q, c]
        arr[i] = DkxStats(int(k[0]), int(k[1]), int(s.lut_size), float(s.dissims_min),
                          float(s.dissims_max), q.ctypes.data_as(C.POINTER(C.c_float)),
                          c.ctypes.data_as(C.POINTER(C.c_float)))
    return arr, len(items), keep


def _collect(out, n):
    res = nlsar_stats_collection()
    for i in range(n):
        o = out[i]
        s = stats(lut_size=o.lut_size)
        s.dissims_min, s.dissims_max = o.dissims_min, o.dissims_max
        s.quantilles = np.ctypeslib.as_array(o.quantilles, shape=(o.lut_size,)).copy()
        s.chi2cdf_inv = np.ctypeslib.as_array(o.chi2cdf_inv, shape=(o.lut_size,)).copy()
        res[params(o.patch_size, o.scale_size)] = s
    return res


def nlsar_train_insar(ampl_master, ampl_slave, phase, patch_sizes, scale_sizes,
                      enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """Train the NL-SAR weighting kernels on a homogeneous area (nlsar.i:45-60)."""
    a1, a2, ph = _f32(ampl_master), _f32(ampl_slave), _f32(phase)
    H, W = _same_shape(a1, a2, ph)
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    n = nP * nS
    out = (DkxStats * max(1, n))()
    store = np.zeros(max(1, n) * 2 * 256, np.float32)
    check(lib().dkx_nlsar_training_insar(a1.ctypes.data, a2.ctypes.data, ph.ctypes.data, H, W, P, nP,
                                         S, nS, out, store.ctypes.data))
    return _collect(out, n)


def _interlace(covmat_raw):
    d, dd, h, w = covmat_raw.shape
    if d != dd:
        raise ValueError("covariance matrices must be square")
    real = np.real(covmat_raw).reshape((-1, h, w))
    imag = np.imag(covmat_raw).reshape((-1, h, w))
    inter = np.empty((2 * d * d, h, w), np.float32)
    inter[::2] = real
    inter[1::2] = imag
    return inter, d, h, w


def nlsar_train(covmat_raw, patch_sizes, scale_sizes, enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """Train on a (d, d, H, W) complex covariance area (nlsar.i:141-154)."""
    inter, d, h, w = _interlace(np.asarray(covmat_raw))
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    n = nP * nS
    out = (DkxStats * max(1, n))()
    store = np.zeros(max(1, n) * 2 * 256, np.float32)
    check(lib().dkx_nlsar_training_covmat(inter.ctypes.data, h, w, d, P, nP, S, nS, out,
                                          store.ctypes.data))
    return _collect(out, n)


def nlsar_insar(ampl_master, ampl_slave, phase, search_window_size, patch_sizes, scale_sizes,
                nlsar_stats, h=15.0, c=49.0, enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """NL-SAR on an interferogram; returns (reflectivity, phase, coherence) (nlsar.i:212-258)."""
    a1, a2, ph = _f32(ampl_master), _f32(ampl_slave), _f32(phase)
    H, W = _same_shape(a1, a2, ph)
    ref = np.zeros((H, W), np.float32)
    phf = np.zeros((H, W), np.float32)
    coh = np.zeros((H, W), np.float32)
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    arr, ns, keep = _stats_structs(nlsar_stats)
    check(lib().dkx_nlsar_insar(a1.ctypes.data, a2.ctypes.data, ph.ctypes.data, ref.ctypes.data,
                                phf.ctypes.data, coh.ctypes.data, H, W, int(search_window_size), P,
                                nP, S, nS, arr, ns, float(h), float(c)))
    return ref, phf, coh


def nlsar(covmat_raw, search_window_size, patch_sizes, scale_sizes, nlsar_stats, h=15.0, c=49.0,
          enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """NL-SAR on (d, d, H, W) complex covariance matrices; returns the filtered matrices in
    the same shape (nlsar.i:157-210)."""
    inter, d, H, W = _interlace(np.asarray(covmat_raw))
    out = np.zeros_like(inter)
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    arr, ns, keep = _stats_structs(nlsar_stats)
    check(lib().dkx_nlsar_covmat(inter.ctypes.data, out.ctypes.data, H, W, d, int(search_window_size),
                                 P, nP, S, nS, arr, ns, float(h), float(c)))
    return out[::2].reshape((d, d, H, W)) + 1j * out[1::2].reshape((d, d, H, W))


def boxcar(ampl_master, ampl_slave, phase, window_width, enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """Boxcar filter; returns (reflectivity, phase, coherence) (boxcar.i:35-62)."""
    a1, a2, ph = _f32(ampl_master), _f32(ampl_slave), _f32(phase)
    H, W = _same_shape(a1, a2, ph)
    ref = np.zeros((H, W), np.float32)
    phf = np.zeros((H, W), np.float32)
    coh = np.zeros((H, W), np.float32)
    check(lib().dkx_boxcar(a1.ctypes.data, a2.ctypes.data, ph.ctypes.data, ref.ctypes.data,
                           phf.ctypes.data, coh.ctypes.data, H, W, int(window_width)))
    return ref, phf, coh


def goldstein(ampl_master, ampl_slave, phase, patch_size, overlap, alpha,
              enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """Goldstein filter; returns (reflectivity, phase, coherence) (goldstein.i:33-75).  As in
    the reference, the coherence is returned as passed in (zeros)."""
    a1, a2, ph = _f32(ampl_master), _f32(ampl_slave), _f32(phase)
    H, W = _same_shape(a1, a2, ph)
    ref = np.zeros((H, W), np.float32)
    phf = np.zeros((H, W), np.float32)
    coh = np.zeros((H, W), np.float32)
    check(lib().dkx_goldstein(a1.ctypes.data, a2.ctypes.data, ph.ctypes.data, ref.ctypes.data,
                              phf.ctypes.data, coh.ctypes.data, H, W, int(patch_size), int(overlap),
                              float(alpha)))
    return ref, phf, coh


def nlinsar(ampl_master, ampl_slave, phase, search_window_size, patch_size, niter, lmin,
            enabled_log_levels=_DEFAULT_LOG_LEVELS):
    """NL-InSAR; returns (reflectivity, phase, coherence) (nlinsar.i:36-74).  The first
    iteration starts from zero estimates, as the reference's Python wrapper passes zeros."""
    a1, a2, ph = _f32(ampl_master), _f32(ampl_slave), _f32(phase)
    H, W = _same_shape(a1, a2, ph)
    ref = np.zeros((H, W), np.float32)
    phf = np.zeros((H, W), np.float32)
    coh = np.zeros((H, W), np.float32)
    check(lib().dkx_nlinsar(a1.ctypes.data, a2.ctypes.data, ph.ctypes.data, ref.ctypes.data,
                            phf.ctypes.data, coh.ctypes.data, H, W, int(search_window_size),
                            int(patch_size), int(niter), int(lmin)))
    return ref, phf, coh


def store_nlsar_stats_collection(nsc, filename):
    """Write the statistics as the reference does: a Boost text archive of
    std::map<nlsar::params, nlsar::stats> (nlsar_training.cpp:99-104)."""
    arr, ns, keep = _stats_structs(nsc)
    check(lib().dkx_nlsar_stats_store(os.fsencode(filename), arr, ns))


def load_nlsar_stats_collection(filename):
    """Read a statistics file written by the reference (Boost text archive) or by
    store_nlsar_stats_collection (nlsar_training.cpp:106-113)."""
    fn = os.fsencode(filename)
    nf = C.c_longlong(0)
    n = lib().dkx_nlsar_stats_load(fn, None, 0, None, 0, C.byref(nf))
    if n < 0:
        check(int(n))
    out = (DkxStats * max(1, n))()
    store = np.zeros(max(1, nf.value), np.float32)
    got = lib().dkx_nlsar_stats_load(fn, out, n, store.ctypes.data, nf.value, C.byref(nf))
    if got < 0:
        check(int(got))
    return _collect(out, n)