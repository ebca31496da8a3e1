"""The autotuner through the C ABI (dkx_nlsar_insar_device_autotune,
dkx_nlinsar_device_autotune, dkx_get_tuned_tile).

Bar: the variant returned is one the parameters allow, every allowed variant has a
positive time and the others -1; the outputs afterwards equal a plain call bit for bit (for
NL-InSAR from the same first-iteration estimates, which the tuner restores before each run);
the record is found again for the same parameters and image size, and a later automatic
call, which uses it, gives the same bits.  Records are cleared afterwards.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PATCHES, SCALES = [3, 5, 7], [1, 3]


@pytest.fixture(scope="module")
def D():
    import despeckcl_amd
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    if hip.device_count() < 1:
        pytest.fail("no HIP device visible")
    despeckcl_amd.lib()
    DV.clear_tuned_tiles()
    yield despeckcl_amd
    DV.clear_tuned_tiles()


@pytest.fixture(scope="module")
def stats(D):
    return D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))


@pytest.mark.parametrize("every_key", [False, True])
def test_nlsar_autotune(D, stats, every_key):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    h, w = 600, 700
    host = S.interferogram(h, w, S.SEED_BASE + 7100)
    inp = [HIP.DeviceArray.from_host(x) for x in host]
    tuned = [HIP.DeviceArray((h, w)) for _ in range(3)]
    plain = [HIP.DeviceArray((h, w)) for _ in range(3)]
    try:
        assert DV.tuned_tile(0, h, w, 21, PATCHES, SCALES, every_key) == -1
        t, ms = DV.nlsar_insar_autotune(*[d.ptr for d in inp], *[d.ptr for d in tuned], h, w, 21, PATCHES, SCALES,
                                        stats, every_key=every_key, reps=1)
        print("variant", t, "ms", ms)
        assert t in (0, 1, 2) and all(m > 0 for m in ms)
        # the fastest variant, or within 2 % of it the one with the fewest rows per thread
        fastest = min(range(3), key=lambda i: ms[i])
        assert t == max(i for i in range(fastest, 3) if ms[i] <= 1.02 * ms[fastest])
        assert DV.tuned_tile(0, h, w, 21, PATCHES, SCALES, every_key) == t
        # another size in the same power-of-two bucket finds the record, another bucket not
        assert DV.tuned_tile(0, h + 10, w, 21, PATCHES, SCALES, every_key) == t
        assert DV.tuned_tile(0, 2 * h, 2 * w, 21, PATCHES, SCALES, every_key) == -1
        after = [d.download() for d in tuned]
        DV.nlsar_insar(*[d.ptr for d in inp], *[d.ptr for d in plain], h, w, 21, PATCHES, SCALES, stats)
        HIP.synchronize()
        for k in range(3):
            assert np.array_equal(after[k], plain[k].download(), equal_nan=True), f"output {k}"
    finally:
        for d in inp + tuned + plain:
            d.free()


@pytest.mark.parametrize("sw,p,lmin,shape,allowed", [(21, 7, 10, (300, 280), (0, 1, 2)),
                                                      (41, 7, 10, (200, 230), (0, 1))])
def test_nlinsar_autotune(D, sw, p, lmin, shape, allowed):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    h, w = shape
    a1, a2, ph = S.interferogram(h, w, S.SEED_BASE + 7200 + sw, n_regions=8, coh_range=(0.3, 0.9))
    init = [(0.5 * (a1 * a1 + a2 * a2) + 1e-3).astype(np.float32), ph.copy(), np.full((h, w), 0.4, np.float32)]
    inp = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    tuned = [HIP.DeviceArray.from_host(x) for x in init]
    plain = [HIP.DeviceArray.from_host(x) for x in init]
    try:
        t, ms = DV.nlinsar_autotune(*[d.ptr for d in inp], *[d.ptr for d in tuned], h, w, sw, p, 2, lmin, reps=1)
        print("shape", t, "ms", ms)
        assert t in allowed
        for k in range(3):
            assert (ms[k] > 0) == (k in allowed), f"shape {k}: {ms[k]}"
        assert DV.tuned_tile(1, h, w, sw, [p]) == t
        after = [d.download() for d in tuned]
        DV.nlinsar(*[d.ptr for d in inp], *[d.ptr for d in plain], h, w, sw, p, 2, lmin)
        HIP.synchronize()
        for k in range(3):
            assert np.array_equal(after[k], plain[k].download(), equal_nan=True), f"output {k}"
        # a later automatic call (which takes the record) from the same estimates
        for d, x in zip(plain, init):
            d.upload(x)
        DV.nlinsar(*[d.ptr for d in inp], *[d.ptr for d in plain], h, w, sw, p, 2, lmin)
        HIP.synchronize()
        for k in range(3):
            assert np.array_equal(after[k], plain[k].download(), equal_nan=True), f"output {k}, tuned call"
    finally:
        for d in inp + tuned + plain:
            d.free()
        D.lib().dkx_release_workspace()
