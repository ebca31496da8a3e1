"""The allocation-free NL-SAR entry point (dkx_nlsar_insar_device_noalloc) against the
allocating one (dkx_nlsar_insar_device), through the C ABI: bit-identical outputs, ENL / alpha
maps and undefined-selection counts, on two calls in a row (the second on a smaller image,
so the scratch array still holds the first call's larger grids), on another stream with
other statistics, and the size limit."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def D():
    import despeckcl_amd
    from despeckcl_amd import hip
    if hip.device_count() < 1:
        pytest.fail("no HIP device visible")
    despeckcl_amd.lib()
    return despeckcl_amd


def _call(D, name, din, dout, H, W, P, S_, stats, enl=None, alp=None, want_ub=False, stream=None):
    from despeckcl_amd import _lib as L
    Pa, nP = L.int_array(P)
    Sa, nS = L.int_array(S_)
    arr, ns, keep = D._stats_structs(stats)
    ub = C.c_ulonglong(0)
    rc = getattr(L.lib(), name)(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, 21, Pa, nP, Sa, nS,
                                arr, ns, 15.0, 49.0, enl.ptr if enl else None, alp.ptr if alp else None,
                                C.byref(ub) if want_ub else None, stream)
    return rc, ub.value


def _run_pair(D, H, W, seed, P, S_, stats, maps, stream=None):
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(H, W, seed, n_regions=6)
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    res = []
    for name in ("dkx_nlsar_insar_device", "dkx_nlsar_insar_device_noalloc"):
        dout = [HIP.DeviceArray((H, W)) for _ in range(3)]
        nk = len(P) * len(S_)
        enl = HIP.DeviceArray((nk, H, W)) if maps else None
        alp = HIP.DeviceArray((nk, H, W)) if maps else None
        rc, ub = _call(D, name, din, dout, H, W, P, S_, stats, enl, alp, want_ub=maps, stream=stream)
        assert rc == 0, D.lib().dkx_last_error()
        HIP.synchronize()
        res.append(([d.download() for d in dout], enl.download() if maps else None,
                    alp.download() if maps else None, ub))
        for d in dout + ([enl, alp] if maps else []):
            d.free()
    for d in din:
        d.free()
    (o1, e1, al1, u1), (o2, e2, al2, u2) = res
    for k in range(3):
        assert np.array_equal(o1[k], o2[k], equal_nan=True), f"{H}x{W} output {k}"
    if maps:
        assert np.array_equal(e1, e2, equal_nan=True) and np.array_equal(al1, al2, equal_nan=True)
        assert u1 == u2
    return o2


def test_noalloc_equals_allocating_entry_twice(D):
    from despeckcl_amd import hip as HIP
    stats = D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))
    P, S_ = [3, 5, 7], [1, 3]
    big = _run_pair(D, 300, 257, 71, P, S_, stats, maps=False)
    small = _run_pair(D, 150, 100, 72, P, S_, stats, maps=True)
    assert np.isfinite(big[0][:-1, :-1]).all() and np.isfinite(small[0][:-1, :-1]).all()
    # a third call on a stream of its own, with a smaller parameter set (other constants)
    st = HIP.Stream()
    sub = D.nlsar_stats_collection()
    for k, v in stats.items():
        if k.patch_size in (3, 5):
            sub[k] = v
    _run_pair(D, 97, 131, 73, [3, 5], S_, sub, maps=True, stream=st.handle)


def test_noalloc_rejects_images_beyond_its_scratch(D):
    from despeckcl_amd import _lib as L
    from despeckcl_amd import hip as HIP
    stats = D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))
    H = W = 2048
    din = [HIP.DeviceArray((1, 1)) for _ in range(3)]
    dout = [HIP.DeviceArray((1, 1)) for _ in range(3)]
    rc, _ = _call(D, "dkx_nlsar_insar_device_noalloc", din, dout, H, W, [3, 5, 7], [1, 3], stats)
    assert rc == L.DKX_ERESOURCE and b"scratch" in D.lib().dkx_last_error()
    for d in din + dout:
        d.free()
