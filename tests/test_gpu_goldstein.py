"""GPU parity of the Goldstein filter (dkx_goldstein[_device], the reference's
despeckcl::goldstein, src/compute/goldstein/goldstein.cpp:39-106 and
goldstein_filter_sub_image.cpp:73-187) against the CPU oracle, through the C ABI.

The oracle transforms exactly (double-precision DFT rounded to float); the device uses a
float radix-2 FFT, as the reference's clFFT does.  Bar, on the complex output
z = ref * exp(i phase), normalised by the RMS of |z| over the image: at most 2e-5 at every
pixel, 99.9th percentile at most 5e-6.  The coherence output is not written (as in the
reference).
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


def _check(g, o, label):
    zg = g[0].astype(np.float64) * np.exp(1j * g[1].astype(np.float64))
    zo = o[0].astype(np.float64) * np.exp(1j * o[1].astype(np.float64))
    rms = np.sqrt(np.mean(np.abs(zo) ** 2))
    e = np.abs(zg - zo) / rms
    q = np.quantile(e, 0.999)
    print(f"{label}: max {e.max():.2e}, q99.9 {q:.2e}")
    assert e.max() <= 2e-5 and q <= 5e-6, label


CASES = [((70, 90), 32, 4, 0.5), ((64, 64), 16, 2, 0.8), ((128, 200), 64, 8, 0.3), ((50, 110), 32, 4, 0.5),
         ((110, 110), 32, 4, 1.0), ((300, 257), 32, 4, 0.5), ((10, 10), 32, 4, 0.5), ((1, 1), 16, 0, 0.5),
         ((33, 7), 16, 4, 0.2)]


@pytest.mark.parametrize("shape,P,ov,alpha", CASES)
def test_goldstein_matches_oracle(D, oracle, shape, P, ov, alpha):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 90 + P + ov)
    g = D.goldstein(a1, a2, ph, P, ov, alpha)
    assert not np.any(g[2])
    _check(g, oracle.goldstein(a1, a2, ph, P, ov, alpha), f"goldstein {shape} P {P} ov {ov} alpha {alpha}")


def test_goldstein_alpha0_adds_raw(D):
    """alpha = 0: the filtered interferogram equals the raw one and the reference adds the two
    (the pack accumulates onto the raw interferogram), so |z| = a1 + a2."""
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(96, 80, 97)
    ref, phase, _ = D.goldstein(a1, a2, ph, 32, 4, 0.0)
    assert np.allclose(ref, a1.astype(np.float64) + a2, rtol=2e-5)
    assert np.abs(np.angle(np.exp(1j * (phase.astype(np.float64) - ph)))).max() <= 1e-4


def test_goldstein_device_entry_equals_host_entry(D):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(200, 170, 98)
    host = D.goldstein(a1, a2, ph, 32, 4, 0.5)
    din = [hip.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [hip.DeviceArray((200, 170)) for _ in range(2)]
    stream = hip.Stream()
    DV.goldstein(*[d.ptr for d in din], *[d.ptr for d in dout], 200, 170, 32, 4, 0.5, stream=stream.handle)
    stream.synchronize()
    for k in range(2):
        assert np.array_equal(dout[k].download(), host[k])
    stream.destroy()
