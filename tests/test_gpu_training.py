"""NL-SAR training on the device (dkx_nlsar_training_insar / _covmat[_device], the
reference's despeckcl::nlsar_training, src/compute/nlsar/training/nlsar_training.cpp:35-96)
against the CPU oracle, through the C ABI.

Bar: bit-identical statistics (dissims_min, dissims_max, the quantile and inverse chi-square
LUTs) for every (patch, scale) key.  Both sides evaluate the reference's host arithmetic
(float determinants, double logs, float sums in pixel order); the quantile LUT is the count
of finite sums below each threshold, an integer on both sides.
"""
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


def _interlace(c):
    d = c.shape[0]
    out = np.empty((2 * d * d,) + c.shape[2:], np.float32)
    out[::2] = c.real.reshape((d * d,) + c.shape[2:])
    out[1::2] = c.imag.reshape((d * d,) + c.shape[2:])
    return out


def _same(got, want, label):
    assert sorted((k.patch_size, k.scale_size) for k in got) == sorted(want.keys()), label
    for k in got:
        g, w = got[k], want[(k.patch_size, k.scale_size)]
        nq = int((g.quantilles != w.quantilles).sum())
        print(f"{label} {tuple(k)}: min {g.dissims_min} / {w.dissims_min}, max {g.dissims_max} / "
              f"{w.dissims_max}, quantiles differing {nq}")
        assert g.lut_size == 256
        assert g.dissims_min == w.dissims_min and g.dissims_max == w.dissims_max, f"{label} {tuple(k)}"
        assert np.array_equal(g.quantilles, w.quantilles), f"{label} {tuple(k)}"
        assert np.array_equal(g.chi2cdf_inv, w.chi2cdf_inv), f"{label} {tuple(k)}"


def test_training_insar_matches_oracle(D, oracle):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.homogeneous_area(32, S.SEED_BASE + 100)
    got = D.nlsar_train_insar(a1, a2, ph, [3, 5, 7], [1, 3])
    _same(got, oracle.nlsar_train_insar(a1, a2, ph, [3, 5, 7], [1, 3]), "insar 32")
    # the committed fixture was written from the same training
    fx = D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))
    _same(got, {(k.patch_size, k.scale_size): v for k, v in fx.items()}, "fixture")


def test_training_example_area(D, oracle):
    """The reference's example: rows/cols 50:75 of its interferogram, patches 3..11,
    scales 1, 3, 5 (examples/insar_test.py:43-56)."""
    z = np.load(os.path.join(GOLDEN, "example_crop.npz"))
    t = np.s_[50:75, 50:75]
    a1, a2, ph = (np.ascontiguousarray(z[k][t]) for k in ("ampl_master", "ampl_slave", "dphase"))
    P, S_ = [3, 5, 7, 9, 11], [1, 3, 5]
    _same(D.nlsar_train_insar(a1, a2, ph, P, S_), oracle.nlsar_train_insar(a1, a2, ph, P, S_), "example")


@pytest.mark.parametrize("dim", [2, 3])
def test_training_covmat_matches_oracle(D, oracle, dim):
    rng = np.random.default_rng(3)
    z = (rng.standard_normal((dim, 20, 20)) + 1j * rng.standard_normal((dim, 20, 20))) / np.sqrt(2)
    cov = np.einsum("ahw,bhw->abhw", z, np.conj(z)).astype(np.complex64)
    got = D.nlsar_train(cov, [3, 5], [1, 3])
    _same(got, oracle.nlsar_train_covmat(_interlace(cov), dim, [3, 5], [1, 3]), f"covmat dim {dim}")


@pytest.mark.parametrize("dim", [2, 3])
def test_training_large_area(D, oracle, dim):
    """A 96x96 area: about half a million patch pairs for patch size 3."""
    from despeckcl_amd import synthetic as S
    c = S.polsar(96, 96, S.SEED_BASE + 300 + dim, dim, n_regions=0)
    got = D.nlsar_train(c, [3, 7], [1, 3])
    _same(got, oracle.nlsar_train_covmat(_interlace(c), dim, [3, 7], [1, 3]), f"large dim {dim}")


def test_training_device_entry_equals_host_entry(D):
    import ctypes as C
    from despeckcl_amd import _lib
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    c = S.polsar(40, 36, 77, 3, n_regions=0)
    host = D.nlsar_train(c, [3, 5], [1, 3])
    inter = _interlace(c)
    d = hip.DeviceArray.from_host(inter)
    out = (_lib.DkxStats * 4)()
    store = np.zeros(4 * 2 * 256, np.float32)
    P, nP = _lib.int_array([3, 5])
    S_, nS = _lib.int_array([1, 3])
    stream = hip.Stream()
    _lib.check(D.lib().dkx_nlsar_training_covmat_device(d.ptr, 40, 36, 3, P, nP, S_, nS, out,
                                                         store.ctypes.data, stream.handle))
    dev = D._collect(out, 4)
    _same(dev, {(k.patch_size, k.scale_size): v for k, v in host.items()}, "device entry")
    stream.destroy()
    del C
