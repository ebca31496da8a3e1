"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (per output pixel: reflectivity and coherence relative error, phase absolute
error in radians, wrapped):

NL-SAR
  stated max error 5e-2 on every pixel whose weighted variance is not rounding noise
  (cond = var / I_nl^2 >= 1e-5, reported by the oracle for the selected key); p99.9 <= 1e-3;
  at most 0.2 % of the pixels (at least 2) above 1e-3.  Pixels with cond < 1e-5 are excluded
  and counted; they must all lie on the last row or column, whose virtual position sees a
  fully replicated neighbourhood (there var is a difference of equal sums, its sign is
  rounding noise, and alpha = 1 - I_nl^2 / var with it, in any implementation).
  Every key's ENL_nobias (relative) and alpha (absolute), on the pixels whose key has
  cond >= 1e-5: max 5e-2, p99 <= 1e-3, at most 1 % above 1e-3 (a LUT-index flip moves the
  weight sums that ENL and alpha are made of, where the output moves less).
NL-InSAR
  one iteration from given estimates: every pixel with an error above 1e-3 must be explained
  by the oracle: either a discrete fragility (its ENL within 1e-4 of lmin; its lmin-smoothing
  threshold tied, within 1e-4 of the next candidate weight, or a tiny weight among tiny ones;
  its weights summing below 1e-18), where a last-bit difference of the weights changes which
  weights are used, or its own sensitivity: equivalent forms of the similarity / KL terms,
  or the same terms in double precision, move its output by at least 1/30 of the error
  (near-identical pixel pairs make the similarity cancel).  At most 1 in 10^4 pixels may
  stay unexplained, each below 1e-2.  Discretely fragile pixels: at most 5 % (images with
  point scatterers, whose weights underflow wholesale in later iterations: stated per
  test); p99 <= 1e-3 on the pixels the oracle flags in no way.  So the stated max error on
  the other pixels is 1e-3.
  Several iterations are checked as a trajectory: the GPU's k-iteration run equals k chained
  single iterations bit for bit, and every one of those iterations passes the check above
  against the oracle run from the GPU's previous estimates.  End to end from zero estimates
  the flips of fragile pixels spread by one search window per iteration, so that comparison
  is distributional (bounds in each test).
The float differences themselves come from the transcendental functions (device library vs
glibc), the order in which the two halves of the search window are summed (NL-SAR), and the
reciprocal / square-root instructions of the NL-InSAR similarity; the LUT truncation (NL-SAR)
and the ENL threshold and smoothing selection (NL-InSAR) turn a few of them into discrete
steps.
At full size the checks are size-independent properties: locality (a crop filtered alone
equals the same region of the full image), tile-shape independence, finite outputs, and the
best-parameter identity.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

COND_MIN = 1e-5          # NL-SAR: var / I_nl^2 below this is rounding noise
NLSAR_MAX_ERR = 5e-2     # stated max error (NL-SAR, well-conditioned pixels)
NLINSAR_MAX_ERR = 1e-3   # stated max error (NL-InSAR, one iteration, non-fragile pixels)


def _errors(g, o):
    ref = np.abs(g[0].astype(np.float64) - o[0]) / np.maximum(np.abs(o[0].astype(np.float64)), 1e-30)
    coh = np.abs(g[2].astype(np.float64) - o[2]) / np.maximum(np.abs(o[2].astype(np.float64)), 1e-30)
    ph = np.abs(np.angle(np.exp(1j * (g[1].astype(np.float64) - o[1].astype(np.float64)))))
    return [np.where(np.isnan(e), np.inf, e) for e in (ref, ph, coh)]


def assert_close(g, o, q=0.999, tol_q=1e-3, flip_frac=0.002, max_tol=None, mask=None, label=""):
    """Quantile q of the error <= tol_q, at most flip_frac of the pixels (at least 2) above
    1e-3, and (max_tol) every error <= max_tol; on the pixels of `mask`."""
    n = g[0].size if mask is None else int(mask.sum())
    allowed = max(2, int(flip_frac * n))
    for name, e in zip(("reflectivity", "phase", "coherence"), _errors(g, o)):
        if mask is not None:
            e = e[mask]
        if e.size == 0:
            continue
        core = np.quantile(e, q)
        nflip = int((e > 1e-3).sum())
        print(f"{label} {name}: max {e.max():.2e} q{q} {core:.2e}  >1e-3: {nflip} of {e.size}")
        assert core <= tol_q, f"{label} {name}: q{q} error {core:.3g}"
        assert nflip <= allowed, f"{label} {name}: {nflip} pixels above 1e-3 (allowed {allowed})"
        if max_tol is not None:
            assert e.max() <= max_tol, f"{label} {name}: max error {e.max():.3g} > {max_tol}"


def _edge_only(mask_excluded):
    """True when every excluded pixel lies on the last row or column."""
    inner = mask_excluded[:-1, :-1]
    return not inner.any()


def nlsar_parity(D, oracle, a1, a2, ph, P, S_, st, label):
    g = D.nlsar_insar(a1, a2, ph, 21, P, S_, st)
    o = oracle.nlsar_insar(a1, a2, ph, 21, P, S_, st, want_cond=True)
    mask = o[4] >= COND_MIN
    excl = ~mask
    ys, xs = np.nonzero(excl)
    print(f"{label}: {int(excl.sum())} pixels with cond < {COND_MIN} excluded:",
          [(int(y), int(x), float(o[4][y, x])) for y, x in zip(ys, xs)][:20])
    assert _edge_only(excl), "an interior pixel has a rounding-noise variance"
    assert excl.sum() <= a1.shape[0] + a1.shape[1]
    for k in range(3):
        assert np.all(np.isfinite(g[k][mask]))
    assert_close(g, o[:3], mask=mask, max_tol=NLSAR_MAX_ERR, label=label)
    return g, o


def _gpu_nlinsar_from(a1, a2, ph, init, sw, p, niter, lmin):
    """NL-InSAR on the GPU starting from given estimates (the C ABI reads the output arrays
    as the first iteration's estimates, like the reference)."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    H, W = a1.shape
    din = [HIP.DeviceArray.from_host(np.ascontiguousarray(x)) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray.from_host(np.ascontiguousarray(x, np.float32)) for x in init]
    DV.nlinsar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, sw, p, niter, lmin)
    out = [d.download() for d in dout]
    for d in din + dout:
        d.free()
    return out


def nlinsar_step(oracle, a1, a2, ph, est, sw, p, lmin, label, frag_frac=0.05):
    """One GPU iteration from `est` against the oracle's from the same estimates.  An error
    above 1e-3 must be explained by the oracle: a discrete fragility of the pixel (ENL at
    lmin, a near-tied or tiny smoothing threshold, underflowing weights), or the pixel's own
    sensitivity (the move of its output under equivalent evaluations) being at least 1/30 of
    the error.  At most 1 in 10^4 pixels may stay unexplained, each below 1e-2.  Returns the
    GPU's estimates."""
    g = _gpu_nlinsar_from(a1, a2, ph, est, sw, p, 1, lmin)
    o = oracle.nlinsar(a1, a2, ph, sw, p, 1, lmin, init=est, want_frag=True)
    flags, sens = o[4], o[5]
    discrete = (flags & (1 | 2 | 4 | 16)) != 0
    e = np.maximum.reduce(_errors(g, o[:3]))
    out = e > NLINSAR_MAX_ERR
    bad = out & ~discrete & ~(e <= 30.0 * sens)
    ys, xs = np.nonzero(bad)
    print(f"{label}: flagged {int((flags != 0).sum())} of {flags.size} (ENL {int(((flags & 1) != 0).sum())}, "
          f"near tie {int(((flags & 2) != 0).sum())}, tiny {int(((flags & 4) != 0).sum())}, "
          f"ill-conditioned {int(((flags & 8) != 0).sum())}, underflow {int(((flags & 16) != 0).sum())}); "
          f"above {NLINSAR_MAX_ERR}: {int(out.sum())}, unexplained: {int(bad.sum())}",
          [(int(y), int(x), float(e[y, x]), float(sens[y, x])) for y, x in zip(ys, xs)][:10])
    assert bad.sum() <= max(1, int(1e-4 * e.size)), f"{label}: {int(bad.sum())} unexplained errors above {NLINSAR_MAX_ERR}"
    assert not bad.any() or e[bad].max() <= 1e-2, f"{label}: unexplained error {e[bad].max():.3g}"
    assert discrete.mean() <= frag_frac, f"{label}: {discrete.mean():.3%} fragile pixels"
    assert np.quantile(e[flags == 0], 0.99) <= 1e-3
    return g


def nlinsar_trajectory(D, oracle, a1, a2, ph, sw, p, niter, lmin, label, frag_frac=0.05):
    """The GPU's niter-iteration run from zero estimates, checked iteration by iteration;
    returns it."""
    est = [np.zeros_like(a1) for _ in range(3)]
    for it in range(niter):
        est = nlinsar_step(oracle, a1, a2, ph, est, sw, p, lmin, f"{label} iteration {it + 1}", frag_frac)
    full = D.nlinsar(a1, a2, ph, sw, p, niter, lmin)
    for k in range(3):
        assert np.array_equal(full[k], est[k], equal_nan=True), f"{label}: {niter} iterations differ from chained single ones"
    return full


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


def _maps_close(g, o, cond, label, rel):
    """Per-key map parity on the pixels whose key has cond >= COND_MIN (relative error when
    rel, else absolute); the bounds of the outputs."""
    for k in range(g.shape[0]):
        m = cond[k] >= COND_MIN
        excl = ~m
        ys, xs = np.nonzero(excl[:-1, :-1])
        if ys.size:
            print(f"{label} key {k}: interior pixels with cond < {COND_MIN}:",
                  [(int(y), int(x), float(cond[k][y, x])) for y, x in zip(ys, xs)])
        assert excl.sum() <= g.shape[1] + g.shape[2]
        d = np.abs(g[k].astype(np.float64) - o[k])
        if rel:
            d = d / np.abs(o[k].astype(np.float64))
        d = np.where(np.isnan(d), np.inf, d)[m]
        n_above = int((d > 1e-3).sum())
        print(f"{label} key {k}: excluded {int(excl.sum())}, max {d.max():.2e}, q99 "
              f"{np.quantile(d, 0.99):.2e}, >1e-3 {n_above} of {d.size}")
        assert np.quantile(d, 0.99) <= 1e-3
        assert n_above <= max(2, int(0.01 * d.size))
        assert d.max() <= NLSAR_MAX_ERR


@pytest.mark.parametrize("shape,seed", [((64, 64), 1), ((97, 41), 4), ((150, 130), 22)])
def test_nlsar_every_key_maps_match_oracle(D, stats, oracle, shape, seed):
    """The bias-corrected ENL and the alpha of every (patch, scale) key, which the
    reference computes and reads back for the selection (nlsar_filter_sub_image.cpp:224-228,
    325-327; compute_enls_nobias.cl:19-39, compute_alphas.cl:19-44), against the oracle's."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S
    H, W = shape
    a1, a2, ph = S.interferogram(H, W, seed)
    din = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [HIP.DeviceArray((H, W)) for _ in range(3)]
    enl = HIP.DeviceArray((6, H, W))
    alp = HIP.DeviceArray((6, H, W))
    ub = DV.nlsar_insar(*[d.ptr for d in din], *[d.ptr for d in dout], H, W, 21, [3, 5, 7], [1, 3],
                        stats, enl_maps=enl.ptr, alpha_maps=alp.ptr, want_ub=True)
    o = oracle.nlsar_insar(a1, a2, ph, 21, [3, 5, 7], [1, 3], stats, want_maps=True)
    assert ub == o[3]
    _maps_close(enl.download(), o[4], o[6], f"ENL_nobias {shape}", rel=True)
    _maps_close(alp.download(), o[5], o[6], f"alpha {shape}", rel=False)


def test_nlsar_statistics_with_extra_keys(D, stats, oracle):
    """A collection with keys beyond patch x scale, as the reference accepts it: when its last
    key is unused the output is key index 0's weighted mean with alpha 0 (the reference's
    default entries, best_params.cpp:27-53, best_alpha_copy.cpp:26-35), against the oracle;
    when the unused keys sort below the last used one the output equals the exact
    collection's bit for bit."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(70, 66, 50)
    P, S_ = [3, 5], [1, 3]
    nlsar_parity(D, oracle, a1, a2, ph, P, S_, stats, "nlsar extra keys, last unused")
    exact = D.nlsar_stats_collection()
    for k, v in stats.items():
        if k.patch_size in (5, 7):
            exact[k] = v
    g = D.nlsar_insar(a1, a2, ph, 21, [5, 7], S_, stats)
    e = D.nlsar_insar(a1, a2, ph, 21, [5, 7], S_, exact)
    for k in range(3):
        assert np.array_equal(g[k], e[k], equal_nan=True)


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
    g = nlinsar_trajectory(D, oracle, a1, a2, ph, 21, 7, niter, 10, f"nlinsar {shape}")
    o = oracle.nlinsar(a1, a2, ph, 21, 7, niter, 10)
    assert_close(g, o[:3], q=0.99, tol_q=1e-3, flip_frac=0.01, label=f"nlinsar {shape} x{niter} end to end")


def test_nlinsar_example_data(D, oracle):
    """The reference example's NL-InSAR parameters (examples/insar_test.py:74-84) on the
    128x128 crop of its example interferogram, 5 iterations, checked as a trajectory.  End to
    end from zero estimates the few fragile pixels' flips spread by a search window per
    iteration (near the replicated image border equal patches meet with opposite argument
    orders, the reference's mirror rule), hence the distributional bound there."""
    z = np.load(os.path.join(GOLDEN, "example_crop.npz"))
    a1, a2, ph = z["ampl_master"], z["ampl_slave"], z["dphase"]
    g = nlinsar_trajectory(D, oracle, a1, a2, ph, 21, 7, 5, 10, "nlinsar example")
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 5, 10)
    assert_close(g, o[:3], q=0.95, tol_q=1e-3, flip_frac=0.05, label="nlinsar example 5 iterations end to end")


def test_nlinsar_patch3(D, oracle):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(60, 60, 8)
    nlinsar_trajectory(D, oracle, a1, a2, ph, 11, 3, 2, 5, "nlinsar p3 sw11")


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
    nlinsar_trajectory(D, oracle, a1, a2, ph, sw, p, 3, lmin, f"nlinsar tiles {shape}", frag_frac=0.6)


def test_nlinsar_large_image_automatic_tile(D):
    """On a large image the automatic choice (18 x 54 blocks of 512 threads) and the 54 x 54
    blocks of 1024 threads give the same bits (1024 x 1024, 2 iterations, point scatterers)."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(1024, 1024, 33, n_regions=32, scatterer_fraction=0.05)
    auto = _nlinsar_with_tile(D, -1, a1, a2, ph, 21, 7, 2, 10)
    t1 = _nlinsar_with_tile(D, 2, a1, a2, ph, 21, 7, 2, 10)
    for k in range(3):
        assert np.array_equal(auto[k], t1[k], equal_nan=True), f"output {k}"


@pytest.mark.parametrize("lmin", [10, 20])
def test_nlinsar_smoothing_ties_and_large_lmin(D, oracle, lmin):
    """Point scatterers make many pixels go through the lmin smoothing, many of them with
    several zero weights tied at the smallest kept weight (those run the exact heap; the
    others the register selection).  lmin = 20 exceeds the register selection's 16 slots,
    so every queued pixel runs the heap.  By the third iteration the smallest kept weight is
    a denormal for a few per cent of the queued pixels; the oracle flags them."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(160, 150, 34, n_regions=16, scatterer_fraction=0.05)
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 3, lmin)
    assert o[3] > 0.3 * 160 * 150, "the case should smooth many pixels"
    g = nlinsar_trajectory(D, oracle, a1, a2, ph, 21, 7, 3, lmin, f"nlinsar smoothing lmin {lmin}", frag_frac=0.6)
    assert_close(g, o[:3], q=0.95, tol_q=1e-3, flip_frac=0.03, label=f"nlinsar smoothing lmin {lmin} x3 end to end")


@pytest.mark.parametrize("sw,p,lmin,shape", [(3, 3, 2, (40, 37)), (5, 7, 4, (33, 50)), (9, 3, 20, (45, 41))])
def test_nlinsar_small_windows(D, oracle, sw, p, lmin, shape):
    """Search windows of 3 to 9 (fewer offsets than one load run of the smoothing, lmin above
    the register selection's 16 slots) against the oracle."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 40 + sw, n_regions=4, scatterer_fraction=0.05)
    nlinsar_trajectory(D, oracle, a1, a2, ph, sw, p, 2, lmin, f"nlinsar sw {sw} p {p}", frag_frac=0.6)


@pytest.mark.parametrize("p,sw,lmin,shape", [(1, 11, 5, (50, 61)), (5, 11, 5, (70, 90)), (9, 15, 10, (64, 70)),
                                             (11, 11, 5, (60, 75)), (13, 11, 5, (58, 66)), (21, 9, 5, (47, 52))])
def test_nlinsar_other_patch_sizes(D, oracle, p, sw, lmin, shape):
    """Patch sizes whose h the reference simulates (quantile_insar, seeded here): the oracle
    gets the filter's own h, three iterations checked as a trajectory."""
    from despeckcl_amd import device as DV
    from despeckcl_amd import synthetic as S
    h = DV.nlinsar_simulate_h(p)
    oracle.set_nlinsar_h_theo(p, h)
    a1, a2, ph = S.interferogram(*shape, S.SEED_BASE + 7300 + p, n_regions=6, coh_range=(0.3, 0.9))
    nlinsar_trajectory(D, oracle, a1, a2, ph, sw, p, 3, lmin, f"nlinsar patch {p} (h_theo {h:.4f})", frag_frac=0.6)


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
