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


def nlsar_covmat(cov, cov_out, height, width, dim, search_window_size, patch_sizes, scale_sizes,
                 nlsar_stats, h=15.0, c=49.0, enl_maps=0, alpha_maps=0, stream=0, want_ub=False):
    """Enqueue NL-SAR on device covariance planes (2*dim*dim interlaced planes in and out).
    Returns the undefined-selection pixel count when want_ub (synchronises), else None."""
    from . import _stats_structs
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    arr, ns, keep = _stats_structs(nlsar_stats)
    ub = C.c_ulonglong(0)
    check(lib().dkx_nlsar_covmat_device(cov, cov_out, int(height), int(width), int(dim),
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


def boxcar(a1, a2, phase, ref, phase_filt, coh, height, width, window_width, stream=0):
    """Enqueue the boxcar filter on device buffers."""
    check(lib().dkx_boxcar_device(a1, a2, phase, ref, phase_filt, coh, int(height), int(width),
                                  int(window_width), stream or None))


def goldstein(a1, a2, phase, ref, phase_filt, height, width, patch_size, overlap, alpha, stream=0):
    """Enqueue the Goldstein filter on device buffers (no coherence output, as the reference)."""
    check(lib().dkx_goldstein_device(a1, a2, phase, ref, phase_filt, None, int(height), int(width),
                                     int(patch_size), int(overlap), float(alpha), stream or None))


def nlinsar_simulate_h(patch_size, seed=5489):
    """The reference's Monte-Carlo h_theo for an NL-InSAR patch size (quantile_insar at alpha
    0.92), with a seeded generator; the filter uses seed 5489 for patch sizes other than 3
    and 7 (dkx_nlinsar_simulate_h)."""
    h = C.c_float(0.0)
    check(lib().dkx_nlinsar_simulate_h(int(patch_size), int(seed), C.byref(h)))
    return h.value


def set_tile(filter_index, tile_index):
    """Block shape of the single-call entry points (dkx_set_tile).  filter 0 (NL-SAR): 0, 1, 2
    = 4, 2, 1 rows per thread; -1 (default) = per pass, 1 row for one patch size, else 2.
    filter 1 (NL-InSAR): 0 = 27 x 54, 1 = 18 x 54, 2 = 54 x 54; -1 (default) = the autotuner's
    record for the parameters and size, else 18 x 54.
    Results do not depend on the setting."""
    check(lib().dkx_set_tile(int(filter_index), int(tile_index)))


def get_tile(filter_index):
    return lib().dkx_get_tile(int(filter_index))


def set_nlinsar_weight_store(mode):
    """Which weights the NL-InSAR weights launches of single calls store for the lmin
    smoothing (dkx_set_nlinsar_weight_store): -1 automatic (default), 0 all, 1 the pixels
    smoothed in the previous iteration and their neighbours, 2 none.  The smoothed pixels'
    missing weights are recomputed, so results do not depend on the setting."""
    check(lib().dkx_set_nlinsar_weight_store(int(mode)))


def get_nlinsar_weight_store():
    return lib().dkx_get_nlinsar_weight_store()


def nlsar_insar_autotune(a1, a2, phase, ref, phase_filt, coh, height, width, search_window_size,
                         patch_sizes, scale_sizes, nlsar_stats, h=15.0, c=49.0, every_key=False, reps=2,
                         stream=0):
    """Time every NL-SAR pass variant on this image (dkx_nlsar_insar_device_autotune), record
    the fastest for its parameters and size; returns (variant, [ms per variant])."""
    from . import _stats_structs
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes)
    arr, ns, keep = _stats_structs(nlsar_stats)
    t = C.c_int(-1)
    ms = (C.c_float * 3)()
    check(lib().dkx_nlsar_insar_device_autotune(a1, a2, phase, ref, phase_filt, coh, int(height), int(width),
                                                int(search_window_size), P, nP, S, nS, arr, ns, float(h),
                                                float(c), 1 if every_key else 0, int(reps), stream or None,
                                                C.byref(t), ms))
    return t.value, [ms[i] for i in range(3)]


def nlinsar_autotune(a1, a2, phase, ref, phase_filt, coh, height, width, search_window_size, patch_size,
                     niter, lmin, reps=1, stream=0):
    """Time every NL-InSAR tile shape on this image (dkx_nlinsar_device_autotune; the
    estimates in ref/phase_filt/coh are restored before each run), record the fastest for
    its parameters and size; returns (shape, [ms per shape, -1 where not allowed])."""
    t = C.c_int(-1)
    ms = (C.c_float * 3)()
    check(lib().dkx_nlinsar_device_autotune(a1, a2, phase, ref, phase_filt, coh, int(height), int(width),
                                            int(search_window_size), int(patch_size), int(niter), int(lmin),
                                            int(reps), stream or None, C.byref(t), ms))
    return t.value, [ms[i] for i in range(3)]


def tuned_tile(filter_index, height, width, search_window_size, patch_sizes, scale_sizes=(), every_key=False):
    """The autotuner's recorded variant for these parameters and image size, -1 if none."""
    P, nP = int_array(patch_sizes)
    S, nS = int_array(scale_sizes) if len(scale_sizes) else (None, 0)
    return lib().dkx_get_tuned_tile(int(filter_index), int(height), int(width), int(search_window_size), P, nP,
                                    S, nS, 1 if every_key else 0)


def clear_tuned_tiles():
    lib().dkx_clear_tuned_tiles()


class NlsarBatch:
    """Batched NL-SAR on device buffers: many images of any size, one parameter set, one
    launch per stage over the tiles of all images (dkx_nlsar_batch_* in the C ABI).

        b = NlsarBatch()
        b.set(images, 21, [3, 5, 7], [1, 3], stats)   # images: (a1, a2, phase, ref, phase_filt, coh, h, w)
        b.set_covmat(images, 3, 21, [3, 5, 7], [1, 3], stats)   # or (covmat, covmat_filt, h, w)
        b.enqueue(stream)          # or b.graph_launch(stream) to capture once and replay
    """

    def __init__(self):
        h = C.c_void_p()
        check(lib().dkx_nlsar_batch_create(C.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if self._h:
            lib().dkx_nlsar_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, images, search_window_size, patch_sizes, scale_sizes, nlsar_stats, h=15.0, c=49.0):
        from . import _stats_structs
        from ._lib import DkxImage
        arr = (DkxImage * len(images))()
        for i, (a1, a2, ph, ref, phf, coh, hh, ww) in enumerate(images):
            arr[i] = DkxImage(a1, a2, ph, ref, phf, coh, int(hh), int(ww))
        P, nP = int_array(patch_sizes)
        S, nS = int_array(scale_sizes)
        st, ns, keep = _stats_structs(nlsar_stats)
        check(lib().dkx_nlsar_batch_set(self._h, arr, len(images), int(search_window_size), P, nP, S, nS,
                                        st, ns, float(h), float(c)))
        self.n = len(images)

    def set_covmat(self, images, dim, search_window_size, patch_sizes, scale_sizes, nlsar_stats, h=15.0, c=49.0):
        """Covariance input of dimension 2 or 3: images (covmat, covmat_filt, h, w), device
        pointers to 2*dim*dim interlaced planes (the layout of nlsar_covmat)."""
        from . import _stats_structs
        from ._lib import DkxCovmatImage
        arr = (DkxCovmatImage * len(images))()
        for i, (cov, out, hh, ww) in enumerate(images):
            arr[i] = DkxCovmatImage(cov, out, int(hh), int(ww))
        P, nP = int_array(patch_sizes)
        S, nS = int_array(scale_sizes)
        st, ns, keep = _stats_structs(nlsar_stats)
        check(lib().dkx_nlsar_batch_set_covmat(self._h, arr, len(images), int(dim), int(search_window_size), P, nP,
                                               S, nS, st, ns, float(h), float(c)))
        self.n = len(images)

    def enqueue(self, stream=0):
        check(lib().dkx_nlsar_batch_enqueue(self._h, stream or None))

    def graph_launch(self, stream):
        check(lib().dkx_nlsar_batch_graph_launch(self._h, stream))

    def set_tile(self, tile):
        check(lib().dkx_nlsar_batch_set_tile(self._h, int(tile)))

    def set_diagnostics(self, on=True):
        """Compute every (patch, scale) key's statistics (needed by ub_pixels)."""
        check(lib().dkx_nlsar_batch_set_diagnostics(self._h, 1 if on else 0))

    def tile(self):
        return lib().dkx_nlsar_batch_get_tile(self._h)

    def autotune(self, stream=0, reps=2):
        t = C.c_int(0)
        ms = (C.c_float * 8)()
        check(lib().dkx_nlsar_batch_autotune(self._h, stream or None, int(reps), C.byref(t), ms))
        return t.value, [ms[i] for i in range(3)]

    def ub_pixels(self):
        out = (C.c_ulonglong * max(1, self.n))()
        check(lib().dkx_nlsar_batch_ub_pixels(self._h, out))
        return [out[i] for i in range(self.n)]


class NlinsarBatch:
    """Batched NL-InSAR on device buffers: many images of any size, one parameter set, one
    launch per stage over all images (dkx_nlinsar_batch_* in the C ABI).  The output arrays
    hold the first iteration's estimates on entry (zero for a plain filter run), as for
    dkx_nlinsar_device.

        b = NlinsarBatch()
        b.set(images, 21, 7, 5, 10)   # images: (a1, a2, phase, ref, phase_filt, coh, h, w)
        b.enqueue(stream)             # or b.graph_launch(stream) to capture once and replay
    """

    def __init__(self):
        h = C.c_void_p()
        check(lib().dkx_nlinsar_batch_create(C.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if self._h:
            lib().dkx_nlinsar_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, images, search_window_size, patch_size, niter, lmin):
        from ._lib import DkxImage
        arr = (DkxImage * len(images))()
        for i, (a1, a2, ph, ref, phf, coh, hh, ww) in enumerate(images):
            arr[i] = DkxImage(a1, a2, ph, ref, phf, coh, int(hh), int(ww))
        check(lib().dkx_nlinsar_batch_set(self._h, arr, len(images), int(search_window_size), int(patch_size),
                                          int(niter), int(lmin)))
        self.n = len(images)

    def enqueue(self, stream=0):
        check(lib().dkx_nlinsar_batch_enqueue(self._h, stream or None))

    def graph_launch(self, stream):
        check(lib().dkx_nlinsar_batch_graph_launch(self._h, stream))

    def smoothed(self):
        """Per image, the pixel-iterations that went through the lmin smoothing (waits)."""
        out = (C.c_ulonglong * max(self.n, 1))()
        check(lib().dkx_nlinsar_batch_smoothed(self._h, out))
        return [int(out[i]) for i in range(self.n)]


class Server:
    """Request loop on the device (dkx_server_* in the C ABI): fixed HBM budget, resident
    NL-SAR LUTs, continuous batching of NL-SAR requests, NL-InSAR requests overlapped on
    streams of their own.  Requests carry device pointers:
    (kind, a1, a2, phase, ref, phase_filt, coh, height, width), kind 0 NL-SAR, 1 NL-InSAR."""

    def __init__(self, hbm_budget_bytes, search_window_size, patch_sizes, scale_sizes, nlsar_stats,
                 h=15.0, c=49.0, nl_search_window_size=21, nl_patch_size=7, nl_niter=5, nl_lmin=10):
        from . import _stats_structs
        P, nP = int_array(patch_sizes)
        S, nS = int_array(scale_sizes)
        st, ns, keep = _stats_structs(nlsar_stats)
        hdl = C.c_void_p()
        check(lib().dkx_server_create(int(hbm_budget_bytes), int(search_window_size), P, nP, S, nS, st, ns,
                                      float(h), float(c), int(nl_search_window_size), int(nl_patch_size),
                                      int(nl_niter), int(nl_lmin), C.byref(hdl)))
        self._h = hdl

    @staticmethod
    def _req(r):
        from ._lib import DkxRequest
        kind, a1, a2, ph, ref, phf, coh, hh, ww = r
        return DkxRequest(int(kind), a1, a2, ph, ref, phf, coh, int(hh), int(ww))

    def submit(self, request):
        i = C.c_longlong(0)
        check(lib().dkx_server_submit(self._h, C.byref(self._req(request)), 0.0, C.byref(i)))
        return i.value

    def wait(self, rid):
        lat = C.c_double(0)
        check(lib().dkx_server_wait(self._h, int(rid), C.byref(lat)))
        return lat.value

    def run_trace(self, requests, arrivals_s):
        """Open-loop replay: request i is submitted arrivals_s[i] seconds after the start.
        Returns (latencies in ms, time to the last completion in s)."""
        from ._lib import DkxRequest
        n = len(requests)
        arr = (DkxRequest * max(1, n))()
        for i, r in enumerate(requests):
            arr[i] = self._req(r)
        t = (C.c_double * max(1, n))(*[float(x) for x in arrivals_s])
        lat = (C.c_double * max(1, n))()
        mk = C.c_double(0)
        check(lib().dkx_server_run_trace(self._h, arr, t, n, lat, C.byref(mk)))
        return [lat[i] for i in range(n)], mk.value

    def counters(self):
        b, im = C.c_longlong(0), C.c_longlong(0)
        check(lib().dkx_server_counters(self._h, C.byref(b), C.byref(im)))
        return b.value, im.value

    def nlinsar_counters(self):
        """(NL-InSAR batches launched, requests they held)."""
        b, im = C.c_longlong(0), C.c_longlong(0)
        check(lib().dkx_server_nlinsar_counters(self._h, C.byref(b), C.byref(im)))
        return b.value, im.value

    def close(self):
        if self._h:
            lib().dkx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
