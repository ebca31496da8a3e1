"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (stated here, applied per output pixel):
  reflectivity and coherence  relative error; phase  absolute error in radians (wrapped)
  NL-SAR     p99.9 <= 1e-3, and at most 0.2 % of the pixels (at least 2) above 1e-3
  NL-InSAR   one iteration from common estimates: p99.9 <= 1e-3, at most 0.5 % above 1e-3;
             up to three iterations end to end: p99 <= 1e-3, at most 1 % above 1e-3;
             five iterations of real data end to end: p90 <= 1e-3, at most 5 % above 1e-3
The float differences come from the transcendental functions (GPU vs glibc), the order in
which the two halves of the search window are summed, and, for NL-InSAR, the argument order
of the mirrored similarities on the reference tiles' borders.  The LUT weights (NL-SAR) and
the ENL threshold and heap ties of the lmin smoothing (NL-InSAR) turn a few of them into
discrete steps, hence the bounded outlier counts.
NL-SAR pixels whose alpha is ill-conditioned are excluded and counted: where the weighted
variance var of the selected parameter is tiny against I_nl^2 (cond = var / I_nl^2 < 0.03,
reported by the oracle), alpha = (var - I_nl^2) / var is rounding noise in any
implementation, the reference's included.  On real images this happens only on the virtual
last row / column, whose neighbourhood is replicated.
At full size (2048^2) the checks are size-independent properties: locality (a crop filtered
alone equals the same region of the full image), tile-shape independence, finite outputs,
and the best-parameter identity.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

COND_MIN = 0.03


def _errors(g, o):
    ref = np.abs(g[0].astype(np.float64) - o[0]) / np.maximum(np.abs(o[0].astype(np.float64)), 1e-30)
    coh = np.abs(g[2].astype(np.float64) - o[2]) / np.maximum(np.abs(o[2].astype(np.float64)), 1e-30)
    ph = np.abs(np.angle(np.exp(1j * (g[1].astype(np.float64) - o[1].astype(np.float64)))))
    return ref, ph, coh


def assert_close(g, o, q=0.999, tol_q=1e-3, flip_frac=0.002, mask=None, label=""):
    n = g[0].size if mask is None else int(mask.sum())
    allowed = max(2, int(flip_frac * n))
    for name, e in zip(("reflectivity", "phase", "coherence"), _errors(g, o)):
        if mask is not None:
            e = e[mask]
        if e.size == 0:
            continue
        e = np.where(np.isnan(e), np.inf, e)
        core = np.quantile(e, q)
        nflip = int((e > 1e-3).sum())
        print(f"{label} {name}: q{q} {core:.2e}  >1e-3: {nflip} of {e.size}")
        assert core <= tol_q, f"{label} {name}: q{q} error {core:.3g}"
        assert nflip <= allowed, f"{label} {name}: {nflip} pixels above 1e-3 (allowed {allowed})"


def nlsar_parity(D, oracle, a1, a2, ph, P, S_, st, label):
    g = D.nlsar_insar(a1, a2, ph, 21, P, S_, st)
    o = oracle.nlsar_insar(a1, a2, ph, 21, P, S_, st, want_cond=True)
    mask = o[4] >= COND_MIN
    print(f"{label}: {int((~mask).sum())} ill-conditioned pixels excluded")
    for k in range(3):
        assert np.all(np.isfinite(g[k][mask]))
    assert_close(g, o[:3], mask=mask, label=label)
    return g, o


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


@pytest.mark.parametrize("shape,seed", [((64, 64), 1), ((50, 77), 2), ((33, 130), 3), ((97, 41), 4)])
def test_nlsar_matches_oracle(D, stats, oracle, shape, seed):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, seed)
    nlsar_parity(D, oracle, a1, a2, ph, [3, 5, 7], [1, 3], stats, f"nlsar {shape}")


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (7, 1), (3, 2), (16, 16)])
def test_nlsar_tiny_images(D, stats, oracle, shape):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 9)
    nlsar_parity(D, oracle, a1, a2, ph, [3, 5, 7], [1, 3], stats, f"nlsar {shape}")


def test_nlsar_example_data_five_patches(D, oracle):
    """The reference's own example (examples/insar_test.py:43-60): patches 3..11, scales 1,3,5,
    statistics trained on rows/cols 50:75, on a 128x128 crop of its example interferogram."""
    z = np.load(os.path.join(GOLDEN, "example_crop.npz"))
    a1, a2, ph = z["ampl_master"], z["ampl_slave"], z["dphase"]
    P, S_ = [3, 5, 7, 9, 11], [1, 3, 5]
    ta = np.s_[50:75, 50:75]
    st = D.nlsar_train_insar(a1[ta], a2[ta], ph[ta], P, S_)
    ost = oracle.nlsar_train_insar(a1[ta], a2[ta], ph[ta], P, S_)
    for k in st:
        assert np.array_equal(st[k].quantilles, ost[(k.patch_size, k.scale_size)].quantilles)
    nlsar_parity(D, oracle, a1, a2, ph, P, S_, st, "nlsar example")


def test_nlsar_selection_identity_on_gpu(D, stats):
    """Every ENL exceeds -1, so the selection keeps the last key (7, 3): the full parameter set
    and the single pair give bit-identical outputs."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(96, 80, 12)
    full = D.nlsar_insar(a1, a2, ph, 21, [3, 5, 7], [1, 3], stats)
    one = D.nlsar_insar(a1, a2, ph, 21, [7], [3], {D.params(7, 3): stats[D.params(7, 3)]})
    for k in range(3):
        assert np.array_equal(full[k], one[k])


def test_nlsar_tile_shapes_bit_identical(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(130, 100, 13)
    outs = []
    old = DV.get_tile(0)
    try:
        for t in (0, 1, 2, -1):   # 4, 2 or 1 rows per thread, automatic per pass
            DV.set_tile(0, t)
            outs.append(D.nlsar_insar(a1, a2, ph, 21, [3, 5, 7], [1, 3], stats))
    finally:
        DV.set_tile(0, old)
    for o in outs[1:]:
        for k in range(3):
            assert np.array_equal(o[k], outs[0][k])


def test_nlsar_device_entry_equals_host_entry(D, stats):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    H, W = 70, 90
    a1, a2, ph = S.interferogram(H, W, 14)
    host = D.nlsar_insar(a1, a2, ph, 21, [3, 5, 7], [1, 3], stats)
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray((H, W)) for _ in range(3)]
    enl = HIP.DeviceArray((6, H, W))
    alp = HIP.DeviceArray((6, H, W))
    ub = DV.nlsar_insar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, 21, [3, 5, 7], [1, 3],
                        stats, enl_maps=enl.ptr, alpha_maps=alp.ptr, want_ub=True)
    assert ub == 0
    for k in range(3):
        assert np.array_equal(dout[k].download(), host[k], equal_nan=True)
    e = enl.download()
    a = alp.download()
    assert np.all(np.isfinite(e)) and np.all(e > 0)
    assert np.all((a >= 0) & np.isfinite(a))


@pytest.mark.parametrize("example", [False, True])
def test_nlsar_last_key_equals_all_keys(D, stats, example):
    """The output depends only on the key the selection returns (the last one,
    best_params.cpp:27-53): the default entry, which computes only that key's statistics,
    gives the same bits as a call that computes every key (requested through the
    undefined-selection count)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    if example:
        z = np.load(os.path.join(GOLDEN, "example_crop.npz"))
        a1, a2, ph = z["ampl_master"], z["ampl_slave"], z["dphase"]
        P, S_ = [3, 5, 7, 9, 11], [1, 3, 5]
        ta = np.s_[50:75, 50:75]
        st = D.nlsar_train_insar(a1[ta], a2[ta], ph[ta], P, S_)
    else:
        a1, a2, ph = S.interferogram(300, 257, 15)
        P, S_, st = [3, 5, 7], [1, 3], stats
    H, W = a1.shape
    host = D.nlsar_insar(a1, a2, ph, 21, P, S_, st)
    din = [HIP.DeviceArray.from_host(np.ascontiguousarray(x)) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray((H, W)) for _ in range(3)]
    ub = DV.nlsar_insar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, 21, P, S_, st, want_ub=True)
    print("undefined-selection pixels", ub)
    for k in range(3):
        assert np.array_equal(dout[k].download(), host[k], equal_nan=True)


def test_nlsar_full_size_locality(D, stats):
    """2048^2 (the benchmark image): finite outputs, no undefined selection, and a crop filtered
    on its own reproduces the full image's interior bit for bit (every output depends only
    on inputs within the 14-pixel overlap, as the reference's tiling assumes)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    H = W = 2048
    a1, a2, ph = S.interferogram(H, W, S.SEED_BASE + 1, n_regions=16, coh_range=(0.2, 0.95))
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray((H, W)) for _ in range(3)]
    ub = DV.nlsar_insar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, 21, [3, 5, 7], [1, 3],
                        stats, want_ub=True)
    full = [d.download() for d in dout]
    assert ub == 0
    # the virtual last row / column can be ill-conditioned (replicated neighbourhood)
    for k in range(3):
        assert np.all(np.isfinite(full[k][:-1, :-1]))
    assert np.all(full[0][:-1, :-1] > 0)
    y0, x0, n, m = 1000, 700, 200, 160
    crop = [np.ascontiguousarray(x[y0:y0 + n, x0:x0 + m]) for x in (a1, a2, ph)]
    c = D.nlsar_insar(*crop, 21, [3, 5, 7], [1, 3], stats)
    o = 15
    for k in range(3):
        assert np.array_equal(c[k][o:n - o, o:m - o], full[k][y0 + o:y0 + n - o, x0 + o:x0 + m - o])


@pytest.mark.parametrize("shape,niter,seed", [((64, 64), 2, 5), ((90, 70), 3, 6), ((55, 120), 1, 7)])
def test_nlinsar_matches_oracle(D, oracle, shape, niter, seed):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, seed)
    g = D.nlinsar(a1, a2, ph, 21, 7, niter, 10)
    o = oracle.nlinsar(a1, a2, ph, 21, 7, niter, 10)
    assert_close(g, o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01, label=f"nlinsar {shape} x{niter}")


def _gpu_nlinsar_from(a1, a2, ph, init, sw, p, niter, lmin):
    """NL-InSAR on the GPU starting from given estimates (the C ABI reads the output arrays
    as the first iteration's estimates, like the reference)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    H, W = a1.shape
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray.from_host(np.ascontiguousarray(x, np.float32)) for x in init]
    DV.nlinsar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, sw, p, niter, lmin)
    return [d.download() for d in dout]


def test_nlinsar_example_data(D, oracle):
    """The reference example's NL-InSAR parameters (examples/insar_test.py:74-84) on the
    128x128 crop of its example interferogram.  Every iteration is checked on its own,
    from the oracle's estimates of the previous iteration; the full 5-iteration run is then
    checked distributionally, because near the replicated image border equal patches are
    evaluated with opposite argument orders (the reference's mirror rule), so their order
    in the lmin smoothing heap depends on the math library's last bits, and such a flip
    spreads by one search window per iteration."""
    z = np.load(os.path.join(GOLDEN, "example_crop.npz"))
    a1, a2, ph = z["ampl_master"], z["ampl_slave"], z["dphase"]
    est = [np.zeros_like(a1) for _ in range(3)]
    for it in range(5):
        o = oracle.nlinsar(a1, a2, ph, 21, 7, 1, 10, init=est)
        g = _gpu_nlinsar_from(a1, a2, ph, est, 21, 7, 1, 10)
        assert_close(g, o[:3], q=0.999, tol_q=1e-3, flip_frac=0.005, label=f"nlinsar example iteration {it + 1}")
        est = list(o[:3])
    g = D.nlinsar(a1, a2, ph, 21, 7, 5, 10)
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 5, 10)
    assert_close(g, o[:3], q=0.9, tol_q=1e-3, flip_frac=0.05, label="nlinsar example 5 iterations")


def test_nlinsar_patch3(D, oracle):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(60, 60, 8)
    g = D.nlinsar(a1, a2, ph, 11, 3, 2, 5)
    o = oracle.nlinsar(a1, a2, ph, 11, 3, 2, 5)
    assert_close(g, o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01, label="nlinsar p3 sw11")


def _nlinsar_with_tile(D, tile, *args):
    from despeckcl_amd import device as DV
    old = DV.get_tile(1)
    try:
        DV.set_tile(1, tile)
        return D.nlinsar(*args)
    finally:
        DV.set_tile(1, old)


@pytest.mark.parametrize("shape,p,sw,lmin", [((300, 257), 7, 21, 10), ((150, 170), 3, 11, 5)])
def test_nlinsar_tile_shapes_bit_identical(D, oracle, shape, p, sw, lmin):
    """The three block shapes (27 x 54 and 18 x 54 with 512 threads, 54 x 54 with 1024) and the
    automatic choice give the same bits: every block emulates the reference's 80 x 80 tiles
    and the per-pixel arithmetic does not depend on the block."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 31, n_regions=12, scatterer_fraction=0.05)
    outs = [_nlinsar_with_tile(D, t, a1, a2, ph, sw, p, 3, lmin) for t in (1, 0, 2, -1)]
    for t, o in zip((0, 2, -1), outs[1:]):
        for k in range(3):
            assert np.array_equal(outs[0][k], o[k], equal_nan=True), f"tile {t}, output {k}"
    o = oracle.nlinsar(a1, a2, ph, sw, p, 3, lmin)
    assert_close(outs[0], o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01, label=f"nlinsar tiles {shape}")


def test_nlinsar_large_image_automatic_tile(D):
    """From 2^20 pixels the automatic choice runs 54 x 54 blocks of 1024 threads; the result
    equals the 18 x 54 blocks' bit for bit (1024 x 1024, 2 iterations, point scatterers)."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(1024, 1024, 33, n_regions=32, scatterer_fraction=0.05)
    auto = _nlinsar_with_tile(D, -1, a1, a2, ph, 21, 7, 2, 10)
    t1 = _nlinsar_with_tile(D, 1, a1, a2, ph, 21, 7, 2, 10)
    for k in range(3):
        assert np.array_equal(auto[k], t1[k], equal_nan=True), f"output {k}"


@pytest.mark.parametrize("lmin", [10, 20])
def test_nlinsar_smoothing_ties_and_large_lmin(D, oracle, lmin):
    """Point scatterers make many pixels go through the lmin smoothing, many of them with
    several zero weights tied at the smallest kept weight (those run the exact heap; the
    others the register selection).  lmin = 20 exceeds the register selection's 16 slots,
    so every queued pixel runs the heap.  Each iteration is checked on its own, from the
    oracle's estimates of the previous one (the oracle's smoothing is the reference's
    std::priority_queue loop).  In the third iteration the smallest kept weight is a
    denormal for about 2.5 % of the queued pixels (362 at lmin 10, 609 at lmin 20, counted
    with the oracle): with so few significant bits, a last-bit difference of exp between the
    device and the host library decides which of those weights are kept, so up to 1 % of
    the pixels may differ.  The 3-iteration run is then checked distributionally, because
    such a flip spreads by one search window per iteration."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(160, 150, 34, n_regions=16, scatterer_fraction=0.05)
    est = [np.zeros_like(a1) for _ in range(3)]
    smoothed = 0
    for it in range(3):
        o = oracle.nlinsar(a1, a2, ph, 21, 7, 1, lmin, init=est)
        g = _gpu_nlinsar_from(a1, a2, ph, est, 21, 7, 1, lmin)
        smoothed += o[3]
        assert_close(g, o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01,
                     label=f"nlinsar smoothing lmin {lmin} iteration {it + 1}")
        est = list(o[:3])
    assert smoothed > 0.3 * 160 * 150, "the case should smooth many pixels"
    g = D.nlinsar(a1, a2, ph, 21, 7, 3, lmin)
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 3, lmin)
    assert_close(g, o[:3], q=0.95, tol_q=1e-3, flip_frac=0.03, label=f"nlinsar smoothing lmin {lmin} x3")


@pytest.mark.parametrize("sw,p,lmin,shape", [(3, 3, 2, (40, 37)), (5, 7, 4, (33, 50)), (9, 3, 20, (45, 41))])
def test_nlinsar_small_windows(D, oracle, sw, p, lmin, shape):
    """Search windows of 3 to 9 (fewer offsets than one load run of the smoothing, lmin above
    the register selection's 16 slots) against the oracle."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 40 + sw, n_regions=4, scatterer_fraction=0.05)
    g = D.nlinsar(a1, a2, ph, sw, p, 2, lmin)
    o = oracle.nlinsar(a1, a2, ph, sw, p, 2, lmin)
    assert_close(g, o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01, label=f"nlinsar sw {sw} p {p}")


def test_nlinsar_large_window_large_image(D):
    """A search window of 41 on a 1024 x 1024 image: the automatic choice must not pick the
    54 x 54 blocks, whose LDS (and reference-tile interior of 34) do not suit this window;
    one iteration, equal to the 18 x 54 blocks' result."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(1024, 1024, 41, n_regions=16)
    auto = _nlinsar_with_tile(D, -1, a1, a2, ph, 41, 7, 1, 10)
    t1 = _nlinsar_with_tile(D, 1, a1, a2, ph, 41, 7, 1, 10)
    for k in range(3):
        assert np.array_equal(auto[k], t1[k], equal_nan=True), f"output {k}"
        assert np.isfinite(auto[k][:-1, :-1]).all()
