"""The oracle's composition pinned end to end.

oracle.c follows the reference stage by stage (tiles, half-window similarity planes and
their launch quadrants, patch-sum planes, the copy_symm mirror, 441-plane weights, the
clamped tile write-back).  oracle/brute.c computes each output pixel straight from what
those stages amount to: clamp-to-edge reads, a direct loop over the full search window, the
pair orientation stated as a rule on positions, and the virtual position the write-back
gives the last row and column.  Both call the same pinned single-pair formulas in the same
summation order, so they must agree bit for bit; an offset, orientation, mirror or
edge-position mistake in either makes them differ.  Reference:
src/compute/nlsar/nlsar_filter_sub_image.cpp:146-390,
cl_wrappers/compute_pixel_similarities_2x2.cpp:51-79, cl_kernels/copy_symm_weights.cl:53-74,
src/compute/nlinsar/nlinsar_sub_image.cpp:22-262, cl_kernels/compute_insar.cl:48-119.

The NL-InSAR closed forms (uniform weights give the window mean, the phase sign, the
keep-previous rule) pin what the brute force and the oracle share.
"""
import numpy as np
import pytest

from despeckcl_amd import synthetic as S

P3, S2 = [3, 5, 7], [1, 3]


@pytest.fixture(scope="module")
def stats(oracle):
    a1, a2, ph = S.homogeneous_area(32, S.SEED_BASE + 100)
    return oracle.nlsar_train_insar(a1, a2, ph, P3, S2)


def _same(a, b):
    """bit-identical, NaN positions included"""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- NL-SAR ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape,seed,sw", [((16, 16), 1, 21), ((9, 13), 2, 21), ((1, 7), 3, 21),
                                           ((6, 1), 4, 21), ((2, 2), 5, 21), ((12, 10), 6, 7)])
def test_nlsar_insar_oracle_equals_brute_force(oracle, stats, shape, seed, sw):
    """Every output pixel, the last row and column included, and every (patch, scale) key's
    bias-corrected ENL and alpha, bit for bit."""
    a1, a2, ph = S.interferogram(*shape, S.SEED_BASE + 300 + seed, n_regions=3)
    o = oracle.nlsar_insar(a1, a2, ph, sw, P3, S2, stats, want_maps=True)
    b = oracle.brute_nlsar_insar(a1, a2, ph, sw, P3, S2, stats)
    for k in range(3):
        assert _same(o[k], b[k]), f"output plane {k}"
    assert _same(o[4], b[3]), "ENL_nobias maps"
    assert _same(o[5], b[4]), "alpha maps"
    assert np.isfinite(b[0]).all()


def test_nlsar_brute_force_detects_a_reversed_orientation(oracle, stats):
    """With every pair's arguments swapped the brute force no longer matches: the check
    above sees the orientation of each dissimilarity, not only its value."""
    a1, a2, ph = S.interferogram(12, 12, S.SEED_BASE + 310, n_regions=3)
    o = oracle.nlsar_insar(a1, a2, ph, 21, P3, S2, stats, want_maps=True)
    b = oracle.brute_nlsar_insar(a1, a2, ph, 21, P3, S2, stats, flip=True)
    assert not _same(o[4], b[3])


def test_nlsar_tiled_oracle_equals_brute_force(oracle, stats):
    """Several reference tiles (the oracle's 40 x 44 tiling): interior pixels do not depend
    on the tiling, and the last row / column take the value of the furthest virtual
    position, which a fully replicated neighbourhood makes the same as overlap + 1 rows
    beyond the edge."""
    a1, a2, ph = S.interferogram(30, 27, S.SEED_BASE + 320, n_regions=4)
    o = oracle.nlsar_insar(a1, a2, ph, 21, P3, S2, stats, tile=(40, 44), want_maps=True)
    b = oracle.brute_nlsar_insar(a1, a2, ph, 21, P3, S2, stats)
    for k in range(3):
        assert _same(o[k], b[k])
    assert _same(o[4], b[3]) and _same(o[5], b[4])


def _interlace(c):
    d = c.shape[0]
    out = np.empty((2 * d * d,) + c.shape[2:], np.float32)
    out[::2] = c.real.reshape((d * d,) + c.shape[2:])
    out[1::2] = c.imag.reshape((d * d,) + c.shape[2:])
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_nlsar_covmat_oracle_equals_brute_force(oracle, dim):
    c = S.polsar(32, 32, S.SEED_BASE + 200 + dim, dim, n_regions=0)
    st = oracle.nlsar_train_covmat(_interlace(c), dim, P3, S2)
    x = _interlace(S.polsar(11, 14, S.SEED_BASE + 330 + dim, dim, n_regions=3))
    o = oracle.nlsar_covmat(x, dim, 11, P3, S2, st, want_maps=True)
    b = oracle.brute_nlsar_covmat(x, dim, 11, P3, S2, st)
    assert _same(o[0], b[0])
    assert _same(o[2], b[1]) and _same(o[3], b[2])


def test_nlsar_superset_statistics(oracle, stats):
    """A collection holding keys beyond patch x scale (a training over [3, 5, 7] used to
    filter with [3, 5]): the reference's enls_nobias map holds a zero-filled entry for the
    unused last key, which every pixel then selects; params2idx and best_alpha_copy read
    default entries for it, so the output is the first key's weighted mean with alpha 0
    (best_params.cpp:27-53, nlsar_filter_sub_image.cpp:224-228, 345-353,
    best_alpha_copy.cpp:26-35).  With the unused keys below the last one the output equals
    the exact collection's."""
    a1, a2, ph = S.interferogram(10, 11, S.SEED_BASE + 340, n_regions=3)
    o = oracle.nlsar_insar(a1, a2, ph, 21, [3, 5], S2, stats)
    b = oracle.brute_nlsar_insar(a1, a2, ph, 21, [3, 5], S2, stats)
    for k in range(3):
        assert _same(o[k], b[k])
    # the first key alone, alpha forced to 0, is what the reference computes here; with
    # alpha 0 the mean is the plain weighted mean, which differs from the exact-collection run
    exact = {k: v for k, v in stats.items() if k[0] in (3, 5)}
    e = oracle.nlsar_insar(a1, a2, ph, 21, [3, 5], S2, exact)
    assert not _same(o[0], e[0])
    # unused keys that sort before the last used one change nothing
    sub = oracle.nlsar_insar(a1, a2, ph, 21, [5, 7], S2, stats)
    exact57 = {k: v for k, v in stats.items() if k[0] in (5, 7)}
    e57 = oracle.nlsar_insar(a1, a2, ph, 21, [5, 7], S2, exact57)
    for k in range(3):
        assert _same(sub[k], e57[k])


def test_nlsar_maps_closed_form(oracle, stats):
    """A constant image: every dissimilarity is 0, so every weight is the LUT weight of
    index 0 (the same w for all 440 offsets) and the centre 1.  Then
    sum w = 1 + 440 w and ENL = (1 + 440 w)^2 / (1 + 440 w^2).  With unit intensities the
    weighted sums of I and I^2 round exactly like sum w, so var = 0 exactly and alpha = 0
    (the (var - I^2) / var quotient is -inf, clamped by max(0, .)); ENL_nobias = ENL.
    (With other intensities var is rounding noise of either sign, and alpha with it: the
    ill-conditioned case the GPU parity tests count separately.)"""
    a1 = np.full((6, 7), 1.0, np.float32)
    a2 = np.full((6, 7), 1.0, np.float32)
    ph = np.full((6, 7), 0.4, np.float32)
    o = oracle.nlsar_insar(a1, a2, ph, 21, P3, S2, stats, want_maps=True)
    enl, alpha = o[4], o[5]
    for si, s in enumerate(S2):
        for pi, p in enumerate(P3):
            st = stats[(p, s)]
            x = float(st.chi2cdf_inv[int(np.float32(st.quantilles[0]) * np.float32(st.lut_size - 1))])
            w = np.exp(-abs(x - 49.0) / 15.0)
            want = (1 + 440 * w) ** 2 / (1 + 440 * w * w)
            k = si * len(P3) + pi
            assert np.allclose(enl[k], want, rtol=1e-4), (p, s)   # float sums of 441 terms
            assert (alpha[k] == 0).all()
    # and the filter returns the input: ref = ((a1 + a2) / 2)^2, phase, coherence 1
    assert np.allclose(o[0], 1.0, rtol=1e-5)
    assert np.allclose(o[1], 0.4, atol=1e-5)
    assert np.allclose(o[2], 1.0, atol=1e-5)


# ---- NL-InSAR -------------------------------------------------------------------------

@pytest.mark.parametrize("shape,p,init,lmin", [((58, 61), 7, "zero", 40), ((58, 61), 7, "prev", 10),
                                               ((31, 40), 3, "prev", 10)])
def test_nlinsar_iteration_oracle_equals_brute_force(oracle, shape, p, init, lmin):
    """One iteration on an image of 2 x 2 reference tiles (80 x 80, step 54 for sw 21 /
    patch 7), from zero estimates (the Python API's first iteration) and from a previous
    estimate: every pixel bit for bit, with the lmin smoothing applied to some of them."""
    a1, a2, ph = S.interferogram(*shape, S.SEED_BASE + 350 + p, n_regions=4)
    est = None
    if init == "prev":
        est = oracle.nlinsar(a1, a2, ph, 21, p, 1, 10)[:3]
    o = oracle.nlinsar(a1, a2, ph, 21, p, 1, lmin, init=est)
    b = oracle.brute_nlinsar(a1, a2, ph, 21, p, 1, lmin, init=est)
    for k in range(3):
        assert _same(o[k], b[k]), f"plane {k}"
    # the lmin smoothing ran for some pixels (the oracle also counts the positions its tiles
    # compute beyond the image edge, so only the brute force's count is per output pixel)
    assert 0 < b[3] < shape[0] * shape[1]


def test_nlinsar_two_iterations_oracle_equals_brute_force(oracle):
    a1, a2, ph = S.interferogram(60, 57, S.SEED_BASE + 360, n_regions=5)
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 2, 10)
    b = oracle.brute_nlinsar(a1, a2, ph, 21, 7, 2, 10)
    for k in range(3):
        assert _same(o[k], b[k])


def test_nlinsar_brute_force_detects_a_reversed_orientation(oracle):
    a1, a2, ph = S.interferogram(58, 61, S.SEED_BASE + 370, n_regions=4)
    est = oracle.nlinsar(a1, a2, ph, 21, 7, 1, 10)[:3]
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 1, 10, init=est)
    b = oracle.brute_nlinsar(a1, a2, ph, 21, 7, 1, 10, init=est, flip=True)
    assert not _same(o[0], b[0])


def test_nlinsar_uniform_weights_give_the_window_mean(oracle):
    """With lmin above the window size every pixel is smoothed; a constant master amplitude
    makes every window position a candidate (a < 2 a), the heap keeps all 441 weights and
    replaces them by their mean (smoothing.cpp:78-129).  compute_insar then returns the
    plain window mean: reflectivity = mean of (a1^2 + a2^2) / 2, coherence =
    |mean of a1 a2 e^{-i phi}| / that mean, phase = arg of it, i.e. minus the input phase
    convention (Q2, precompute_filter_values.cl:36-37, compute_insar.cl:103)."""
    H, W, sw = 30, 33, 21
    rng = np.random.default_rng(5)
    a1 = np.full((H, W), 2.0, np.float32)
    a2 = rng.uniform(0.5, 3.0, (H, W)).astype(np.float32)
    ph = rng.uniform(-0.6, 1.4, (H, W)).astype(np.float32)
    o = oracle.nlinsar(a1, a2, ph, sw, 7, 1, 500)
    # one 80 x 80 tile, all 54 x 54 of its interior positions smoothed
    assert o[3] == 54 * 54
    wsh = sw // 2
    # virtual positions of the last row / column: the end of the last tile's interior
    def virt(n):
        step = 80 - 2 * 13
        v = np.arange(n)
        v[-1] = ((n - 1 + 13) // step) * step + step - 1
        return v
    vy, vx = virt(H), virt(W)
    a = 0.5 * (a1.astype(np.float64) ** 2 + a2.astype(np.float64) ** 2)
    x = a1.astype(np.float64) * a2 * np.exp(-1j * ph.astype(np.float64))
    ref = np.zeros((H, W))
    z = np.zeros((H, W), np.complex128)
    for i in range(H):
        ys = np.clip(np.arange(vy[i] - wsh, vy[i] + wsh + 1), 0, H - 1)
        for j in range(W):
            xs = np.clip(np.arange(vx[j] - wsh, vx[j] + wsh + 1), 0, W - 1)
            ref[i, j] = a[np.ix_(ys, xs)].mean()
            z[i, j] = x[np.ix_(ys, xs)].mean()
    assert np.allclose(o[0], ref, rtol=2e-5)
    assert np.allclose(o[2], np.abs(z) / ref, rtol=2e-5, atol=2e-6)
    dphi = np.angle(np.exp(1j * (o[1] - np.angle(z))))
    assert np.abs(dphi).max() < 2e-5
    # Q2: a constant input phase comes back negated
    o2 = oracle.nlinsar(a1, a2, np.full((H, W), 0.7, np.float32), sw, 7, 1, 500)
    assert np.allclose(o2[1], -0.7, atol=1e-6)


def test_nlinsar_keeps_the_previous_estimate(oracle):
    """Q5 (compute_insar.cl:37-46, 106-117): where the new reflectivity, phase or coherence
    is 0, NaN or inf the previous estimate stays.  Zero amplitudes give a = x = 0 in every
    window that lies inside the zero region: reflectivity 0 / N = 0, phase atan2(0, 0) = 0,
    coherence 0 / 0 = NaN."""
    H, W = 40, 44
    a1, a2, ph = S.interferogram(H, W, S.SEED_BASE + 380, n_regions=3)
    a1 = a1.copy()
    a2 = a2.copy()
    a1[:, :30] = 0
    a2[:, :30] = 0
    init = [np.full((H, W), 5.0, np.float32), np.full((H, W), 0.3, np.float32),
            np.full((H, W), 0.5, np.float32)]
    o = oracle.nlinsar(a1, a2, ph, 21, 7, 1, 10, init=init)
    inside = np.s_[:, :30 - 10]   # windows (sw 21) entirely within the zero columns
    assert (o[0][inside] == 5.0).all() and (o[1][inside] == np.float32(0.3)).all()
    assert (o[2][inside] == np.float32(0.5)).all()
    # away from the zero region new values are written
    assert not (o[0][:, 40:] == 5.0).any()
    b = oracle.brute_nlinsar(a1, a2, ph, 21, 7, 1, 10, init=init)
    for k in range(3):
        assert _same(o[k], b[k])
