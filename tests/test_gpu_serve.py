"""The request loop (dkx_server_*) through the C ABI: mixed NL-SAR / NL-InSAR requests of
uneven size, admitted against an HBM budget, NL-SAR requests batched, NL-InSAR requests on
their own streams.

Bar: every request's result is bit-identical (NaN positions included) to the same filter run
alone through the single-call device entry point; a request larger than the budget fails
without blocking the others.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PATCHES, SCALES = [3, 5, 7], [1, 3]
NITER, LMIN = 2, 10


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


class Req:
    def __init__(self, hip, S, kind, h, w, seed):
        self.kind, self.h, self.w = kind, h, w
        host = S.interferogram(h, w, seed)
        self.inp = [hip.DeviceArray.from_host(x) for x in host]
        self.out = [hip.DeviceArray((h, w)) for _ in range(3)]

    def tuple(self):
        return (self.kind, *[d.ptr for d in self.inp], *[d.ptr for d in self.out], self.h, self.w)

    def clear(self):
        for d in self.out:
            d.zero()

    def result(self):
        return [d.download() for d in self.out]


def _alone(D, r, stats, stream):
    from despeckcl_amd import device as DV
    r.clear()
    args = [*[d.ptr for d in r.inp], *[d.ptr for d in r.out], r.h, r.w]
    if r.kind == 0:
        DV.nlsar_insar(*args, 21, PATCHES, SCALES, stats, stream=stream)
    else:
        DV.nlinsar(*args, 21, 7, NITER, LMIN, stream=stream)


def _same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


SHAPES = [(0, 64, 64), (1, 60, 70), (0, 33, 130), (0, 97, 41), (1, 90, 45), (0, 128, 128),
          (1, 33, 33), (0, 7, 200), (0, 150, 60), (1, 81, 100), (0, 20, 20), (0, 300, 90)]


def test_served_results_equal_alone(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    assert DV.get_tile(0) == -1, "the per-call entry points run the automatic pass variant"
    stream = hip.Stream()
    reqs = [Req(hip, S, k, h, w, 500 + i) for i, (k, h, w) in enumerate(SHAPES)]
    alone = []
    for r in reqs:
        _alone(D, r, stats, stream.handle)
        stream.synchronize()
        alone.append(r.result())
    srv = DV.Server(2 << 30, 21, PATCHES, SCALES, stats, nl_niter=NITER, nl_lmin=LMIN)
    for label, arrivals in (("burst", [0.0] * len(reqs)), ("staggered", [0.002 * i for i in range(len(reqs))])):
        for r in reqs:
            r.clear()
        hip.synchronize()
        lat, span = srv.run_trace([r.tuple() for r in reqs], arrivals)
        print(label, "latencies ms", [round(x, 2) for x in lat], "makespan", round(span, 4))
        assert all(x > 0 for x in lat)
        for i, r in enumerate(reqs):
            assert _same(r.result(), alone[i]), f"{label}: request {i} {SHAPES[i]} differs from its run alone"
    batches, images = srv.counters()
    print("batches", batches, "batched images", images)
    assert batches >= 1 and images == 2 * sum(1 for k, _, _ in SHAPES if k == 0)
    srv.close()
    stream.destroy()


def test_request_over_budget_fails_alone(D, stats):
    from despeckcl_amd import _lib
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    small = Req(hip, S, 0, 48, 48, 610)
    big = Req(hip, S, 1, 600, 600, 611)      # NL-InSAR workspace of about 1.3 GB
    srv = DV.Server(64 << 20, 21, PATCHES, SCALES, stats, nl_niter=NITER, nl_lmin=LMIN)
    a = srv.submit(small.tuple())
    b = srv.submit(big.tuple())
    assert srv.wait(a) > 0
    with pytest.raises(_lib.DespeckleError):
        srv.wait(b)
    # the server keeps serving
    c = srv.submit(small.tuple())
    assert srv.wait(c) > 0
    srv.close()
