"""Batched NL-SAR through the C ABI (dkx_nlsar_batch_*): one launch per stage over the tiles
of images of uneven size.

Bar: bit-identical to the single-image device entry point with the same pass variant, for
every image of the batch (NaN positions included), whatever the other images are; HIP
graph replays, interleaved with eager work, equal the eager run every time.  Parity of the
single-image path against the oracle is covered in test_gpu_parity.py.
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
    from despeckcl_amd import hip
    if hip.device_count() < 1:
        pytest.fail("no HIP device visible")
    despeckcl_amd.lib()
    return despeckcl_amd


@pytest.fixture(scope="module")
def stats(D):
    return D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))


class Img:
    """Device buffers of one interferogram and its outputs."""

    def __init__(self, hip, S, h, w, seed):
        self.h, self.w = h, w
        self.host = S.interferogram(h, w, seed)
        self.inp = [hip.DeviceArray.from_host(x) for x in self.host]
        self.out = [hip.DeviceArray((h, w)) for _ in range(3)]

    def tuple(self):
        return (*[d.ptr for d in self.inp], *[d.ptr for d in self.out], self.h, self.w)

    def clear(self):
        for d in self.out:
            d.zero()

    def result(self):
        return [d.download() for d in self.out]


def _single(D, img, stats, stream):
    from despeckcl_amd import device as DV
    DV.nlsar_insar(*img.tuple()[:8], 21, PATCHES, SCALES, stats, stream=stream)


def _same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


SHAPES = [(64, 64), (50, 77), (130, 33), (97, 41), (1, 9), (7, 1), (300, 180), (40, 257)]


@pytest.mark.parametrize("tile", [0, 1, 2])
def test_batch_equals_single_calls(D, stats, tile):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    stream = hip.Stream()
    imgs = [Img(hip, S, h, w, 300 + i) for i, (h, w) in enumerate(SHAPES)]
    old = DV.get_tile(0)
    try:
        DV.set_tile(0, tile)
        singles = []
        for im in imgs:
            im.clear()
            _single(D, im, stats, stream.handle)
            stream.synchronize()
            singles.append(im.result())
        b = DV.NlsarBatch()
        b.set([im.tuple() for im in imgs], 21, PATCHES, SCALES, stats)
        b.set_tile(tile)
        assert b.tile() == tile
        # without and with the statistics of every key (the diagnostics): same images
        for diag in (False, True):
            b.set_diagnostics(diag)
            for im in imgs:
                im.clear()
            hip.synchronize()
            b.enqueue(stream.handle)
            stream.synchronize()
            for i, im in enumerate(imgs):
                print(f"tile {tile} diagnostics {diag} image {SHAPES[i]}")
                assert _same(im.result(), singles[i]), f"image {i} {SHAPES[i]} differs from its single call"
        assert b.ub_pixels() == [0] * len(imgs)
        b.set_diagnostics(False)
        with pytest.raises(ValueError, match="diagnostics"):
            b.ub_pixels()
        b.close()
    finally:
        DV.set_tile(0, old)
        stream.destroy()


def test_batch_result_independent_of_companions(D, stats):
    """The same image gives the same result in two different batches (alone and in a mix)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    stream = hip.Stream()
    a = Img(hip, S, 120, 90, 777)
    others = [Img(hip, S, h, w, 800 + i) for i, (h, w) in enumerate([(33, 300), (200, 64), (5, 5)])]
    b1 = DV.NlsarBatch()
    b1.set([a.tuple()], 21, PATCHES, SCALES, stats)
    b1.enqueue(stream.handle)
    stream.synchronize()
    alone = a.result()
    a.clear()
    b2 = DV.NlsarBatch()
    b2.set([others[0].tuple(), a.tuple(), others[1].tuple(), others[2].tuple()], 21, PATCHES, SCALES, stats)
    b2.set_tile(b1.tile())
    b2.enqueue(stream.handle)
    stream.synchronize()
    assert _same(a.result(), alone)
    b1.close()
    b2.close()
    stream.destroy()


def test_batch_graph_replays_equal_eager(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    stream = hip.Stream()
    imgs = [Img(hip, S, h, w, 900 + i) for i, (h, w) in enumerate([(80, 70), (33, 129), (150, 40)])]
    b = DV.NlsarBatch()
    b.set([im.tuple() for im in imgs], 21, PATCHES, SCALES, stats)
    b.enqueue(stream.handle)
    stream.synchronize()
    eager = [im.result() for im in imgs]
    extra = Img(hip, S, 64, 64, 999)
    for rep in range(4):
        for im in imgs:
            im.clear()
        hip.synchronize()
        b.graph_launch(stream.handle)
        stream.synchronize()
        for i, im in enumerate(imgs):
            assert _same(im.result(), eager[i]), f"replay {rep} image {i} differs from the eager run"
        # the program's own eager work between replays
        _single(D, extra, stats, stream.handle)
        stream.synchronize()
    b.close()
    stream.destroy()


def test_batch_autotune_keeps_results(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    stream = hip.Stream()
    imgs = [Img(hip, S, h, w, 1100 + i) for i, (h, w) in enumerate([(256, 256), (100, 300)])]
    b = DV.NlsarBatch()
    b.set([im.tuple() for im in imgs], 21, PATCHES, SCALES, stats)
    t, ms = b.autotune(stream.handle, reps=1)
    print("autotune", t, ms)
    assert t in (0, 1, 2) and b.tile() == t
    assert all(m > 0 for m in ms)
    b.enqueue(stream.handle)
    stream.synchronize()
    got = [im.result() for im in imgs]
    b.set_tile(t)
    for im in imgs:
        im.clear()
    hip.synchronize()
    b.enqueue(stream.handle)
    stream.synchronize()
    for i, im in enumerate(imgs):
        assert _same(im.result(), got[i])
    b.close()
    stream.destroy()


def test_batch_rejects_bad_arguments(D, stats):
    from despeckcl_amd import device as DV
    b = DV.NlsarBatch()
    with pytest.raises(ValueError):
        b.set([], 21, PATCHES, SCALES, stats)
    with pytest.raises(ValueError):
        b.set([(0, 0, 0, 0, 0, 0, 8, 8)], 21, PATCHES, SCALES, stats)
    with pytest.raises(ValueError):
        b.enqueue(0)
    b.close()
