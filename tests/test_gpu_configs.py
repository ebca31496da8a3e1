"""The benchmark workloads themselves, on the GPU, at their full sizes.

  NL-InSAR 256 x 256, 8 regions, 5 iterations, lmin 10: against the oracle, iteration by
    iteration and end to end.
  NL-InSAR 4096 x 4096, 32 regions and 5 % point scatterers, 10 iterations: finite outputs,
    locality (a crop on the reference's tile grid filtered alone equals the full image beyond
    13 pixels per iteration from the crop's edges), and a 512 x 512 crop against the oracle
    iteration by iteration.
  Batched NL-SAR, the seeded 64 images with sides in [256, 2048]: every image bit for bit
    equal to its single call, eager and replayed from the HIP graph.
  Batched NL-InSAR, ten images of uneven size: every image bit for bit equal to its single
    call, eager and replayed from the HIP graph.
  Served stream, the seeded 2000 requests (70 % NL-SAR / 30 % NL-InSAR, sides in
    [128, 4096]) all at once: 100 seeded requests bit for bit equal to the same request
    filtered alone.
Tolerances as stated in test_gpu_parity.py.
"""
import os

import numpy as np
import pytest

from test_gpu_parity import assert_close, nlinsar_step, nlinsar_trajectory

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


@pytest.fixture(scope="module")
def stats(D):
    return D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))


def test_nlinsar_256_five_iterations(D, oracle):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(256, 256, S.SEED_BASE + 0, n_regions=8, coh_range=(0.3, 0.9))
    g = nlinsar_trajectory(D, oracle, a1, a2, ph, 21, 7, 5, 10, "nlinsar 256^2")
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 5, 10)
    # end to end from zero estimates: the fragile pixels' flips spread by one search window
    # per iteration
    assert_close(g, o[:3], q=0.99, tol_q=5e-3, flip_frac=0.02, label="nlinsar 256^2 x5 end to end")


def test_nlinsar_4096_ten_iterations_with_scatterers(D, oracle):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    n, niter = 4096, 10
    a1, a2, ph = S.interferogram(n, n, S.SEED_BASE + 2, n_regions=32, coh_range=(0.3, 0.9),
                                 scatterer_fraction=0.05)
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray((n, n)) for _ in range(3)]
    for d in dout:
        d.zero()
    nsm = DV.nlinsar(*[d.ptr for d in din], *[d.ptr for d in dout], n, n, 21, 7, niter, 10,
                     want_count=True)
    full = [d.download() for d in dout]
    for d in din + dout:
        d.free()
    D.lib().dkx_release_workspace()   # the 4096^2 workspace (about 60 GB)
    print("smoothed pixel-iterations", nsm)
    assert nsm > 0
    for k in range(3):
        assert np.isfinite(full[k]).all()
    assert (full[0] > 0).all() and (full[2] >= 0).all() and (full[2] <= 1).all()
    # locality: each iteration reads estimates within 13 pixels (search window 21, patch 7);
    # a crop whose origin is a multiple of the reference's tile step (54) keeps its tile grid,
    # so the pairs' argument order, and the bits, are the same
    y0, x0, m = 54 * 30, 54 * 41, 540
    crop = [np.ascontiguousarray(x[y0:y0 + m, x0:x0 + m]) for x in (a1, a2, ph)]
    c = D.nlinsar(*crop, 21, 7, niter, 10)
    r = 13 * niter + 1
    for k in range(3):
        assert np.array_equal(c[k][r:m - r, r:m - r], full[k][y0 + r:y0 + m - r, x0 + r:x0 + m - r],
                              equal_nan=True), f"output {k}"
    # a 512 x 512 crop of the same data against the oracle, iteration by iteration
    s = [np.ascontiguousarray(x[1000:1512, 2000:2512]) for x in (a1, a2, ph)]
    nlinsar_trajectory(D, oracle, *s, 21, 7, niter, 10, "nlinsar 512^2 crop with scatterers", frag_frac=0.6)
    D.lib().dkx_release_workspace()


def test_batched_64_images_equal_single_calls(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    shapes = S.tile_shapes(64, S.SEED_BASE + 3)
    imgs = []
    for i, (h, w) in enumerate(shapes):
        host = S.interferogram(h, w, S.SEED_BASE + 3000 + i, n_regions=16, coh_range=(0.2, 0.95))
        imgs.append(([HIP.DeviceArray.from_host(x) for x in host], [HIP.DeviceArray((h, w)) for _ in range(3)],
                     [HIP.DeviceArray((h, w)) for _ in range(3)], h, w))
    print("pixels", sum(h * w for *_x, h, w in imgs))
    stream = HIP.Stream()
    b = DV.NlsarBatch()
    try:
        b.set([(*[d.ptr for d in inp], *[d.ptr for d in out], h, w) for inp, out, _s, h, w in imgs],
              21, [3, 5, 7], [1, 3], stats)
        b.enqueue(stream.handle)
        stream.synchronize()
        eager = [[d.download() for d in out] for _i, out, _s, _h, _w in imgs]
        b.graph_launch(stream.handle)
        b.graph_launch(stream.handle)
        stream.synchronize()
        for i, (inp, out, single, h, w) in enumerate(imgs):
            DV.nlsar_insar(*[d.ptr for d in inp], *[d.ptr for d in single], h, w, 21, [3, 5, 7], [1, 3],
                           stats, stream=stream.handle)
            stream.synchronize()
            for k in range(3):
                one = single[k].download()
                assert np.array_equal(eager[i][k], one, equal_nan=True), f"image {i} ({h}x{w}) output {k}"
                assert np.array_equal(out[k].download(), one, equal_nan=True), f"graph, image {i} output {k}"
    finally:
        b.close()
        for inp, out, single, _h, _w in imgs:
            for d in inp + out + single:
                d.free()


NLI_BATCH_SHAPES = [(1, 7), (5, 3), (64, 64), (80, 80), (81, 200), (300, 257), (512, 512), (97, 1000),
                    (700, 130), (256, 256)]


@pytest.mark.parametrize("sw,p,lmin,niter,tile", [(21, 7, 10, 3, -1), (21, 7, 10, 2, 2), (11, 3, 4, 2, -1)])
def test_nlinsar_batch_equals_single_calls(D, sw, p, lmin, niter, tile):
    """Ten images of uneven size (one row, a few pixels, up to 512 x 512), half of them from
    nonzero first-iteration estimates: every image bit for bit equal to its single call, eager
    and on two replays of the HIP graph; the smoothing counts too."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    old = DV.get_tile(1)
    DV.set_tile(1, tile)
    imgs = []
    try:
        for i, (h, w) in enumerate(NLI_BATCH_SHAPES):
            a1, a2, ph = S.interferogram(h, w, S.SEED_BASE + 5000 + i, n_regions=8, coh_range=(0.2, 0.95))
            if i % 2:
                init = [(0.5 * (a1 * a1 + a2 * a2) + 1e-3).astype(np.float32), ph.copy(),
                        np.full((h, w), 0.3, np.float32)]
            else:
                init = [np.zeros((h, w), np.float32) for _ in range(3)]
            imgs.append(([HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)],
                         [HIP.DeviceArray.from_host(x) for x in init],
                         [HIP.DeviceArray.from_host(x) for x in init], init, h, w))
        stream = HIP.Stream()
        b = DV.NlinsarBatch()
        try:
            b.set([(*[d.ptr for d in inp], *[d.ptr for d in out], h, w) for inp, out, _s, _i, h, w in imgs],
                  sw, p, niter, lmin)
            b.enqueue(stream.handle)
            stream.synchronize()
            eager = [[d.download() for d in out] for _i, out, _s, _n, _h, _w in imgs]
            counts = b.smoothed()
            print("smoothed pixel-iterations", counts)
            for i, (inp, _o, single, _n, h, w) in enumerate(imgs):
                n_single = DV.nlinsar(*[d.ptr for d in inp], *[d.ptr for d in single], h, w, sw, p, niter,
                                      lmin, stream=stream.handle, want_count=True)
                assert counts[i] == n_single, f"image {i} ({h}x{w}) smoothing count"
                for k in range(3):
                    assert np.array_equal(eager[i][k], single[k].download(), equal_nan=True), \
                        f"image {i} ({h}x{w}) output {k}"
            for replay in range(2):
                for _inp, out, _s, init, _h, _w in imgs:
                    for d, x in zip(out, init):
                        d.upload(x)
                HIP.synchronize()
                b.graph_launch(stream.handle)
                stream.synchronize()
                for i, (_inp, out, _s, _n, h, w) in enumerate(imgs):
                    for k in range(3):
                        assert np.array_equal(out[k].download(), eager[i][k], equal_nan=True), \
                            f"graph replay {replay}, image {i} ({h}x{w}) output {k}"
        finally:
            b.close()
    finally:
        DV.set_tile(1, old)
        for inp, out, single, _n, _h, _w in imgs:
            for d in inp + out + single:
                d.free()
        D.lib().dkx_release_workspace()


def test_served_stream_2000_requests(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    n = 2000
    kinds, shapes, origins, _arr = S.served_stream(n)
    scenes = []
    for k in range(4):
        host = S.served_scene(k)
        scenes.append([HIP.DeviceArray.from_host(x) for x in host])
        del host
    reqs, bufs = [], []
    for i in range(n):
        h, w = shapes[i]
        y0, x0 = origins[i]
        inp = [HIP.DeviceArray((h, w)) for _ in range(3)]
        for d, src in zip(inp, scenes[i % 4]):
            HIP.copy2d(d.ptr, w * 4, src.ptr + (y0 * 4096 + x0) * 4, 4096 * 4, w * 4, h)
        out = [HIP.DeviceArray((h, w)) for _ in range(3)]
        bufs.append((inp, out))
        reqs.append((int(kinds[i]), *[d.ptr for d in inp], *[d.ptr for d in out], h, w))
    HIP.synchronize()
    D.lib().dkx_release_workspace()
    try:
        # 78 GB of request buffers, a 100 GiB budget (the largest NL-InSAR request needs 60 GB)
        srv = DV.Server(100 << 30, 21, [3, 5, 7], [1, 3], stats, 15.0, 49.0, 21, 7, 5, 10)
        try:
            lat, span = srv.run_trace(reqs, [0.0] * n)
            batches, batched = srv.counters()
            nbatches, nbatched = srv.nlinsar_counters()
        finally:
            srv.close()
        print(f"2000 requests in {span:.2f} s, {batches} NL-SAR batches of {batched} images, "
              f"{nbatches} NL-InSAR batches of {nbatched}")
        assert min(lat) >= 0
        assert batched == int((kinds == 0).sum()) and batches < batched
        assert nbatched == int((kinds == 1).sum()) and nbatches < nbatched
        rng = np.random.default_rng(2000)
        sample = sorted(rng.choice(n, 100, replace=False).tolist())
        assert {int(kinds[i]) for i in sample} == {0, 1}
        _check_alone(D, DV, HIP, stats, reqs, bufs, sample)
    finally:
        for inp, out in bufs:
            for d in inp + out:
                d.free()
        for sc in scenes:
            for d in sc:
                d.free()
        D.lib().dkx_release_workspace()


def _check_alone(D, DV, HIP, stats, reqs, bufs, sample):
    stream = HIP.Stream()
    for i in sample:
        k, a1, a2, ph, r1, r2, r3, h, w = reqs[i]
        alone = [HIP.DeviceArray((h, w)) for _ in range(3)]
        if k == 0:
            DV.nlsar_insar(a1, a2, ph, *[d.ptr for d in alone], h, w, 21, [3, 5, 7], [1, 3], stats,
                           stream=stream.handle)
        else:
            for d in alone:
                d.zero()
            DV.nlinsar(a1, a2, ph, *[d.ptr for d in alone], h, w, 21, 7, 5, 10, stream=stream.handle)
        stream.synchronize()
        for j in range(3):
            assert np.array_equal(bufs[i][1][j].download(), alone[j].download(), equal_nan=True), \
                f"request {i} (kind {k}, {h}x{w}) output {j}"
        for d in alone:
            d.free()
