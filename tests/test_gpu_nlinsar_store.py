"""NL-InSAR weight store modes (dkx_set_nlinsar_weight_store), through the C ABI.

The lmin smoothing (smoothing.cpp:39-76 in the reference) needs the search-window weights of
the pixels whose ENL falls below lmin.  The weights launches store every weight (mode 0),
only those of the pixels the previous iteration smoothed and their neighbours (mode 1), or
none (mode 2); a queued pixel whose weights were not stored has them recomputed by
nlinsar_fill_weights with the weights kernel's operations in its order.  Every mode, and the
automatic one, must give the same bits and the same count of smoothed pixels; mode 2 sends
every queued pixel through the recomputation, mode 1 a mix of stored and recomputed ones.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import despeckcl_amd
    from despeckcl_amd import hip
    if hip.device_count() < 1:
        pytest.fail("no HIP device visible")
    despeckcl_amd.lib()
    return despeckcl_amd


def _run(D, mode, tile, a1, a2, ph, sw, p, niter, lmin, init=None):
    """The device entry point with the given store mode and tile shape; returns the three
    outputs and the number of smoothed pixel-iterations."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    H, W = a1.shape
    old_mode, old_tile = DV.get_nlinsar_weight_store(), DV.get_tile(1)
    d_in = [HIP.DeviceArray.from_host(np.ascontiguousarray(x, np.float32)) for x in (a1, a2, ph)]
    d_out = [HIP.DeviceArray((H, W)) for _ in range(3)]
    try:
        DV.set_nlinsar_weight_store(mode)
        DV.set_tile(1, tile)
        for k, d in enumerate(d_out):
            if init is None:
                d.zero()
            else:
                d.upload(init[k])
        n = DV.nlinsar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], H, W, sw, p, niter, lmin,
                       want_count=True)
        HIP.synchronize()
        return [d.download() for d in d_out], n
    finally:
        DV.set_nlinsar_weight_store(old_mode)
        DV.set_tile(1, old_tile)
        for d in d_in + d_out:
            d.free()


def _same(ref, other, label):
    (ro, rn), (oo, on) = ref, other
    assert rn == on, f"{label}: smoothed pixel-iterations {on} vs {rn}"
    for k in range(3):
        assert np.array_equal(ro[k], oo[k], equal_nan=True), \
            f"{label}: output {k} differs at {np.count_nonzero(ro[k] != oo[k])} pixels"


CASES = [
    # (shape, seed, sw, p, niter, lmin, tile, nonzero initial estimates)
    ((256, 256), 41, 21, 7, 10, 10, 1, False),
    ((300, 257), 42, 21, 7, 6, 10, 0, True),
    ((160, 150), 43, 21, 3, 5, 20, 2, False),
    ((130, 120), 44, 15, 5, 4, 8, 1, True),
    ((97, 111), 45, 11, 9, 3, 5, 0, False),
    ((1, 9), 46, 5, 3, 3, 25, 1, False),
]


@pytest.mark.parametrize("shape,seed,sw,p,niter,lmin,tile,init", CASES)
def test_store_modes_bit_identical(D, shape, seed, sw, p, niter, lmin, tile, init):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, seed, n_regions=12, scatterer_fraction=0.05)
    est = None
    if init:
        # a first pass as the initial estimates: the KL term is on from the first iteration
        est, _ = _run(D, 0, tile, a1, a2, ph, sw, p, 1, lmin)
    base = _run(D, 0, tile, a1, a2, ph, sw, p, niter, lmin, est)
    assert base[1] > 0, "the case should smooth some pixels"
    for mode in (2, 1, -1):
        _same(base, _run(D, mode, tile, a1, a2, ph, sw, p, niter, lmin, est), f"mode {mode}")


def test_store_modes_scatterer_image_1024(D):
    """A 1024^2 crop of the 4096^2 scatterer workload, 10 iterations, the automatic tile
    shape (54 x 54): the automatic mode stores all weights in iterations 2-3 and predicts
    from iteration 4 on."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(4096, 4096, S.SEED_BASE + 2, n_regions=32, coh_range=(0.3, 0.9),
                                 scatterer_fraction=0.05)
    a1, a2, ph = (np.ascontiguousarray(x[:1024, :1024]) for x in (a1, a2, ph))
    base = _run(D, 0, -1, a1, a2, ph, 21, 7, 10, 10)
    assert base[1] > 0.1 * 1024 * 1024
    _same(base, _run(D, -1, -1, a1, a2, ph, 21, 7, 10, 10), "automatic")
    _same(base, _run(D, 2, -1, a1, a2, ph, 21, 7, 10, 10), "none stored")


def test_store_mode_rejected(D):
    from despeckcl_amd import device as DV
    with pytest.raises(Exception):
        DV.set_nlinsar_weight_store(3)
    assert DV.get_nlinsar_weight_store() == -1


BATCH_SHAPES = [(1, 7), (5, 3), (64, 64), (81, 200), (300, 257), (200, 130), (256, 256), (97, 400)]


def test_batch_store_modes_bit_identical(D):
    """Batches under every store mode (the mode in force when the batch is set): the
    predicted iterations' marks, the dense decision and the recomputation run over all images
    at once, the thin edge blocks' stored pixels go to compact slots too.  Every image equals
    its single call with all weights stored, eagerly and on two replays of the HIP graph
    (marks and recomputation captured)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    sw, p, niter, lmin = 21, 7, 6, 10
    old = DV.get_nlinsar_weight_store()
    imgs = []
    try:
        for i, (h, w) in enumerate(BATCH_SHAPES):
            a1, a2, ph = S.interferogram(h, w, 7100 + i, n_regions=6, scatterer_fraction=0.05)
            imgs.append(([HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)],
                         [HIP.DeviceArray((h, w)) for _ in range(3)], h, w))
        stream = HIP.Stream()
        ref = []
        DV.set_nlinsar_weight_store(0)
        for inp, out, h, w in imgs:
            for d in out:
                d.zero()
            n = DV.nlinsar(*[d.ptr for d in inp], *[d.ptr for d in out], h, w, sw, p, niter, lmin,
                           stream=stream.handle, want_count=True)
            ref.append(([d.download() for d in out], n))
        assert sum(r[1] for r in ref) > 0
        for mode in (0, 1, 2, -1):
            DV.set_nlinsar_weight_store(mode)
            b = DV.NlinsarBatch()
            try:
                b.set([(*[d.ptr for d in inp], *[d.ptr for d in out], h, w) for inp, out, h, w in imgs],
                      sw, p, niter, lmin)
                for run in range(3):
                    for _inp, out, _h, _w in imgs:
                        for d in out:
                            d.zero()
                    HIP.synchronize()
                    if run == 0:
                        b.enqueue(stream.handle)
                    else:
                        b.graph_launch(stream.handle)
                    stream.synchronize()
                    counts = b.smoothed()
                    for i, (_inp, out, h, w) in enumerate(imgs):
                        if run == 0:
                            assert counts[i] == ref[i][1], f"mode {mode}, image {i} ({h}x{w}) smoothing count"
                        for k in range(3):
                            assert np.array_equal(out[k].download(), ref[i][0][k], equal_nan=True), \
                                f"mode {mode}, run {run}, image {i} ({h}x{w}) output {k}"
            finally:
                b.close()
    finally:
        DV.set_nlinsar_weight_store(old)
        for inp, out, _h, _w in imgs:
            for d in inp + out:
                d.free()
        D.lib().dkx_release_workspace()
