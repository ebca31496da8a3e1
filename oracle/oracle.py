"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, always as the checker.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class OrcStats(C.Structure):
    _fields_ = [("patch_size", C.c_int), ("scale_size", C.c_int), ("lut_size", C.c_int),
                ("dissims_min", C.c_float), ("dissims_max", C.c_float),
                ("quantilles", C.POINTER(C.c_float)), ("chi2cdf_inv", C.POINTER(C.c_float))]


def build(force=False):
    """Compile liboracle.so with the committed Makefile (gcc; a no-op when it is newer than
    its sources)."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_chi2_inv.restype = C.c_double
        L.orc_chi2_inv.argtypes = [C.c_double, C.c_double]
        L.orc_chi2_cdf.restype = C.c_double
        L.orc_chi2_cdf.argtypes = [C.c_double, C.c_double]
        for name in ("orc_det_covmat_2x2",):
            getattr(L, name).restype = C.c_float
            getattr(L, name).argtypes = [C.c_float] * 4
        L.orc_det_covmat_3x3.restype = C.c_float
        L.orc_det_covmat_3x3.argtypes = [C.c_float] * 9
        L.orc_pixel_similarity_2x2.restype = C.c_float
        L.orc_pixel_similarity_2x2.argtypes = [_f32p, _f32p, C.c_int]
        L.orc_pixel_similarity_3x3.restype = C.c_float
        L.orc_pixel_similarity_3x3.argtypes = [_f32p, _f32p, C.c_int]
        L.orc_enl_nobias.restype = C.c_float
        L.orc_enl_nobias.argtypes = [C.c_float] * 3
        L.orc_nlinsar_pixel_similarity.restype = C.c_float
        L.orc_nlinsar_pixel_similarity.argtypes = [C.c_float] * 6
        L.orc_nlinsar_pixel_kl.restype = C.c_float
        L.orc_nlinsar_pixel_kl.argtypes = [C.c_float] * 6
        L.orc_training_dissims.restype = C.c_longlong
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _stats_array(stats):
    """stats: mapping (patch, scale) -> object with lut_size, dissims_min, dissims_max,
    quantilles, chi2cdf_inv.  Returns (ctypes array, keepalive list)."""
    keep = []
    arr = (OrcStats * len(stats))()
    for i, ((p, s), st) in enumerate(sorted(stats.items())):
        q = _f32(st.quantilles)
        c = _f32(st.chi2cdf_inv)
        keep += [q, c]
        arr[i] = OrcStats(p, s, int(st.lut_size), float(st.dissims_min), float(st.dissims_max),
                          q.ctypes.data_as(C.POINTER(C.c_float)),
                          c.ctypes.data_as(C.POINTER(C.c_float)))
    return arr, keep


def _ints(v):
    a = (C.c_int * len(v))(*[int(x) for x in v])
    return a


class _Stats:
    __slots__ = ("lut_size", "dissims_min", "dissims_max", "quantilles", "chi2cdf_inv")

    def __init__(self, lut_size, dmin, dmax, q, c):
        self.lut_size, self.dissims_min, self.dissims_max = lut_size, dmin, dmax
        self.quantilles, self.chi2cdf_inv = q, c


def chi2_inv(p, nu=49.0):
    return lib().orc_chi2_inv(float(p), float(nu))


def chi2_cdf(x, nu=49.0):
    return lib().orc_chi2_cdf(float(x), float(nu))


def stats_build(dissims, lut_size=256):
    d = _f32(dissims)
    q = np.zeros(lut_size, np.float32)
    c = np.zeros(lut_size, np.float32)
    dmin, dmax = C.c_float(), C.c_float()
    m = lib().orc_stats_build(d.ctypes.data_as(C.c_void_p), C.c_int(d.size), C.c_int(lut_size),
                              C.byref(dmin), C.byref(dmax), q.ctypes.data_as(C.c_void_p),
                              c.ctypes.data_as(C.c_void_p))
    return _Stats(lut_size, dmin.value, dmax.value, q, c), m


def _vp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def nlsar_insar(a1, a2, phi, search_window_size, patch_sizes, scale_sizes, stats,
                h=15.0, c=49.0, tile=(0, 0), init=None, want_cond=False, want_maps=False):
    """Returns (ref, phase, coh, ub_pixels), plus the conditioning map when want_cond
    (per pixel min over the diagonal of var / I_nl^2 of the selected parameter), plus
    (enl_maps, alpha_maps, cond_maps) of shape (n_patch*n_scale, H, W), key index
    scale_index*n_patch + patch_index, when want_maps."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    out = [np.zeros((H, W), np.float32) for _ in range(3)] if init is None else [_f32(x).copy() for x in init]
    cond = np.zeros((H, W), np.float32) if want_cond else None
    nk = len(patch_sizes) * len(scale_sizes)
    enl = np.zeros((nk, H, W), np.float32) if want_maps else None
    alp = np.zeros((nk, H, W), np.float32) if want_maps else None
    cnd = np.zeros((nk, H, W), np.float32) if want_maps else None
    arr, keep = _stats_array(stats)
    ub = C.c_longlong(0)
    rc = lib().orc_nlsar_insar_maps(
        _vp(a1), _vp(a2), _vp(phi), _vp(out[0]), _vp(out[1]), _vp(out[2]), _vp(cond), _vp(enl),
        _vp(alp), _vp(cnd), C.c_int(H), C.c_int(W), C.c_int(search_window_size),
        _ints(patch_sizes), C.c_int(len(patch_sizes)), _ints(scale_sizes), C.c_int(len(scale_sizes)),
        arr, C.c_int(len(stats)), C.c_float(h), C.c_float(c), C.c_int(tile[0]), C.c_int(tile[1]),
        C.byref(ub))
    if rc != 0:
        raise ValueError(f"oracle nlsar_insar failed ({rc})")
    res = [out[0], out[1], out[2], ub.value]
    if want_cond:
        res.append(cond)
    if want_maps:
        res += [enl, alp, cnd]
    return tuple(res)


def nlsar_covmat(covmat, dim, search_window_size, patch_sizes, scale_sizes, stats,
                 h=15.0, c=49.0, tile=(0, 0), want_cond=False, want_maps=False):
    """covmat: (2*dim*dim, H, W) interlaced planes.  Returns (filtered planes, ub_pixels),
    plus the conditioning map (as nlsar_insar) when want_cond, plus (enl_maps, alpha_maps,
    cond_maps) when want_maps."""
    cov = _f32(covmat)
    H, W = cov.shape[-2:]
    out = np.zeros_like(cov)
    cond = np.zeros((H, W), np.float32) if want_cond else None
    nk = len(patch_sizes) * len(scale_sizes)
    enl = np.zeros((nk, H, W), np.float32) if want_maps else None
    alp = np.zeros((nk, H, W), np.float32) if want_maps else None
    cnd = np.zeros((nk, H, W), np.float32) if want_maps else None
    arr, keep = _stats_array(stats)
    ub = C.c_longlong(0)
    rc = lib().orc_nlsar_covmat_maps(
        _vp(cov), _vp(out), _vp(cond), _vp(enl), _vp(alp), _vp(cnd), C.c_int(H), C.c_int(W),
        C.c_int(dim), C.c_int(search_window_size), _ints(patch_sizes), C.c_int(len(patch_sizes)),
        _ints(scale_sizes), C.c_int(len(scale_sizes)), arr, C.c_int(len(stats)), C.c_float(h),
        C.c_float(c), C.c_int(tile[0]), C.c_int(tile[1]), C.byref(ub))
    if rc != 0:
        raise ValueError(f"oracle nlsar_covmat failed ({rc})")
    res = [out, ub.value]
    if want_cond:
        res.append(cond)
    if want_maps:
        res += [enl, alp, cnd]
    return tuple(res)


def brute_nlsar_insar(a1, a2, phi, search_window_size, patch_sizes, scale_sizes, stats,
                      h=15.0, c=49.0, flip=False):
    """Per-pixel brute force (brute.c).  Returns (ref, phase, coh, enl_maps, alpha_maps)."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    out = [np.zeros((H, W), np.float32) for _ in range(3)]
    nk = len(patch_sizes) * len(scale_sizes)
    enl = np.zeros((nk, H, W), np.float32)
    alp = np.zeros((nk, H, W), np.float32)
    arr, keep = _stats_array(stats)
    rc = lib().brute_nlsar_insar(
        _vp(a1), _vp(a2), _vp(phi), _vp(out[0]), _vp(out[1]), _vp(out[2]), C.c_int(H), C.c_int(W),
        C.c_int(search_window_size), _ints(patch_sizes), C.c_int(len(patch_sizes)),
        _ints(scale_sizes), C.c_int(len(scale_sizes)), arr, C.c_int(len(stats)), C.c_float(h),
        C.c_float(c), _vp(enl), _vp(alp), C.c_int(1 if flip else 0))
    if rc != 0:
        raise ValueError(f"brute nlsar_insar failed ({rc})")
    return out[0], out[1], out[2], enl, alp


def brute_nlsar_covmat(covmat, dim, search_window_size, patch_sizes, scale_sizes, stats,
                       h=15.0, c=49.0, flip=False):
    """Per-pixel brute force (brute.c).  Returns (filtered planes, enl_maps, alpha_maps)."""
    cov = _f32(covmat)
    H, W = cov.shape[-2:]
    out = np.zeros_like(cov)
    nk = len(patch_sizes) * len(scale_sizes)
    enl = np.zeros((nk, H, W), np.float32)
    alp = np.zeros((nk, H, W), np.float32)
    arr, keep = _stats_array(stats)
    rc = lib().brute_nlsar_covmat(
        _vp(cov), _vp(out), C.c_int(H), C.c_int(W), C.c_int(dim), C.c_int(search_window_size),
        _ints(patch_sizes), C.c_int(len(patch_sizes)), _ints(scale_sizes), C.c_int(len(scale_sizes)),
        arr, C.c_int(len(stats)), C.c_float(h), C.c_float(c), _vp(enl), _vp(alp),
        C.c_int(1 if flip else 0))
    if rc != 0:
        raise ValueError(f"brute nlsar_covmat failed ({rc})")
    return out, enl, alp


def brute_nlinsar(a1, a2, phi, search_window_size, patch_size, niter, lmin, init=None, tile=80,
                  flip=False):
    """NL-InSAR by the per-pixel brute force (brute.c), niter iterations.  Returns
    (ref, phase, coh, n_smoothed)."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    est = [np.zeros((H, W), np.float32) for _ in range(3)] if init is None else [_f32(x).copy() for x in init]
    nsm_total = 0
    for _ in range(niter):
        new = [np.zeros((H, W), np.float32) for _ in range(3)]
        ns = C.c_longlong(0)
        rc = lib().brute_nlinsar_iteration(
            _vp(a1), _vp(a2), _vp(phi), _vp(est[0]), _vp(est[1]), _vp(est[2]), _vp(new[0]),
            _vp(new[1]), _vp(new[2]), C.c_int(H), C.c_int(W), C.c_int(search_window_size),
            C.c_int(patch_size), C.c_int(lmin), C.c_int(tile), C.byref(ns), C.c_int(1 if flip else 0))
        if rc != 0:
            raise ValueError(f"brute nlinsar failed ({rc})")
        est = new
        nsm_total += ns.value
    return est[0], est[1], est[2], nsm_total


def boxcar(a1, a2, phi, window_width, tile=0):
    """Returns (ref, phase, coh)."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    out = [np.zeros((H, W), np.float32) for _ in range(3)]
    rc = lib().orc_boxcar(a1.ctypes.data_as(C.c_void_p), a2.ctypes.data_as(C.c_void_p),
                          phi.ctypes.data_as(C.c_void_p), *[o.ctypes.data_as(C.c_void_p) for o in out],
                          C.c_int(H), C.c_int(W), C.c_int(window_width), C.c_int(tile))
    if rc != 0:
        raise ValueError(f"oracle boxcar failed ({rc})")
    return tuple(out)


def goldstein(a1, a2, phi, patch_size, overlap, alpha):
    """Returns (ref, phase); the reference leaves coh_filt untouched."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    ref = np.zeros((H, W), np.float32)
    ph = np.zeros((H, W), np.float32)
    rc = lib().orc_goldstein(a1.ctypes.data_as(C.c_void_p), a2.ctypes.data_as(C.c_void_p),
                             phi.ctypes.data_as(C.c_void_p), ref.ctypes.data_as(C.c_void_p),
                             ph.ctypes.data_as(C.c_void_p), C.c_int(H), C.c_int(W),
                             C.c_int(patch_size), C.c_int(overlap), C.c_float(alpha))
    if rc != 0:
        raise ValueError(f"oracle goldstein failed ({rc})")
    return ref, ph


def _training_out(n, stats_arr, qbuf, cbuf):
    res = {}
    for i in range(n):
        s = stats_arr[i]
        res[(s.patch_size, s.scale_size)] = _Stats(s.lut_size, s.dissims_min, s.dissims_max,
                                                   qbuf[i].copy(), cbuf[i].copy())
    return res


def nlsar_train_insar(a1, a2, phi, patch_sizes, scale_sizes):
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    n = len(patch_sizes) * len(scale_sizes)
    arr = (OrcStats * n)()
    qbuf = np.zeros((n, 256), np.float32)
    cbuf = np.zeros((n, 256), np.float32)
    rc = lib().orc_nlsar_training_insar(
        a1.ctypes.data_as(C.c_void_p), a2.ctypes.data_as(C.c_void_p), phi.ctypes.data_as(C.c_void_p),
        C.c_int(H), C.c_int(W), _ints(patch_sizes), C.c_int(len(patch_sizes)), _ints(scale_sizes),
        C.c_int(len(scale_sizes)), arr, qbuf.ctypes.data_as(C.c_void_p), cbuf.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError(f"oracle training failed ({rc})")
    return _training_out(n, arr, qbuf, cbuf)


def nlsar_train_covmat(covmat, dim, patch_sizes, scale_sizes):
    cov = _f32(covmat)
    H, W = cov.shape[-2:]
    n = len(patch_sizes) * len(scale_sizes)
    arr = (OrcStats * n)()
    qbuf = np.zeros((n, 256), np.float32)
    cbuf = np.zeros((n, 256), np.float32)
    rc = lib().orc_nlsar_training_covmat(
        cov.ctypes.data_as(C.c_void_p), C.c_int(H), C.c_int(W), C.c_int(dim), _ints(patch_sizes),
        C.c_int(len(patch_sizes)), _ints(scale_sizes), C.c_int(len(scale_sizes)), arr,
        qbuf.ctypes.data_as(C.c_void_p), cbuf.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError(f"oracle training failed ({rc})")
    return _training_out(n, arr, qbuf, cbuf)


def training_dissims(covmat, dim, patch_size, scale_size):
    cov = _f32(covmat)
    H, W = cov.shape[-2:]
    h_ov, w_ov = H - scale_size + 1, W - scale_size + 1
    npatch = -(-h_ov // patch_size) * -(-w_ov // patch_size)
    cap = npatch * (npatch - 1) // 2
    out = np.zeros(max(cap, 1), np.float32)
    n = lib().orc_training_dissims(cov.ctypes.data_as(C.c_void_p), C.c_int(H), C.c_int(W),
                                   C.c_int(dim), C.c_int(patch_size), C.c_int(scale_size),
                                   out.ctypes.data_as(C.c_void_p), C.c_longlong(cap))
    return out[:n]


def set_nlinsar_h_theo(patch_size, h_theo):
    """h_theo for a patch size other than 3 and 7 (the reference draws it at random)."""
    if lib().orc_nlinsar_set_h_theo(C.c_int(patch_size), C.c_float(h_theo)) != 0:
        raise ValueError("invalid patch size")


def nlinsar(a1, a2, phi, search_window_size, patch_size, niter, lmin, init=None, tile=80,
            want_frag=False):
    """Returns (ref, phase, coh, n_smoothed), plus the last iteration's fragility flags
    (uint8) and output sensitivity (see orc_nlinsar_frag) when want_frag."""
    a1, a2, phi = _f32(a1), _f32(a2), _f32(phi)
    H, W = a1.shape
    out = [np.zeros((H, W), np.float32) for _ in range(3)] if init is None else [_f32(x).copy() for x in init]
    frag = np.zeros((H, W), np.float32) if want_frag else None
    sens = np.zeros((H, W), np.float32) if want_frag else None
    ns = C.c_longlong(0)
    rc = lib().orc_nlinsar_frag(
        _vp(a1), _vp(a2), _vp(phi), _vp(out[0]), _vp(out[1]), _vp(out[2]), _vp(frag), _vp(sens),
        C.c_int(H), C.c_int(W), C.c_int(search_window_size),
        C.c_int(patch_size), C.c_int(niter), C.c_int(lmin), C.c_int(tile), C.byref(ns))
    if rc != 0:
        raise ValueError(f"oracle nlinsar failed ({rc})")
    if want_frag:
        return out[0], out[1], out[2], ns.value, frag.astype(np.uint8), sens
    return out[0], out[1], out[2], ns.value


def call(name, *args):
    """Call a per-kernel restatement with numpy arrays / ints."""
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(a.ctypes.data_as(C.c_void_p))
        elif isinstance(a, float):
            conv.append(C.c_float(a))
        else:
            conv.append(C.c_int(int(a)))
    return getattr(lib(), name)(*conv)
