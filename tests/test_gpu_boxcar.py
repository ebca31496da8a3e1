"""GPU parity of the boxcar filter (dkx_boxcar[_device], the reference's despeckcl::boxcar,
src/compute/boxcar/boxcar.cpp:35-92 and cl_kernels/boxcar_kernel.cl:19-69) against the CPU
oracle, through the C ABI.

Bar: the reflectivity is bit-identical (the same float additions in the same order and a
correctly rounded division on both sides).  Phase and coherence go through cos / sin /
atan2, whose last bits differ between the device library and glibc: the complex coherence
coh * exp(i phase) agrees to 1e-5 in absolute value at every pixel.  The last row and
column (the reference's tile-edge values) are included.
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
    assert np.array_equal(g[0], o[0]), f"{label}: reflectivity differs ({int((g[0] != o[0]).sum())} pixels)"
    zg = g[2].astype(np.float64) * np.exp(1j * g[1].astype(np.float64))
    zo = o[2].astype(np.float64) * np.exp(1j * o[1].astype(np.float64))
    e = np.abs(zg - zo)
    print(f"{label}: complex coherence max error {e.max():.2e}")
    assert e.max() <= 1e-5, label


@pytest.mark.parametrize("shape,w", [((64, 64), 5), ((100, 130), 3), ((600, 530), 7), ((37, 513), 15),
                                     ((1030, 70), 1), ((129, 65), 9), ((1, 1), 5), ((7, 1), 3), ((1, 9), 15),
                                     ((2, 2), 11), ((16, 64), 13)])
def test_boxcar_matches_oracle(D, oracle, shape, w):
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(*shape, 70 + w)
    _check(D.boxcar(a1, a2, ph, w), oracle.boxcar(a1, a2, ph, w), f"boxcar {shape} w {w}")


def test_boxcar_device_entry_equals_host_entry(D):
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.interferogram(300, 257, 79)
    host = D.boxcar(a1, a2, ph, 7)
    din = [hip.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    dout = [hip.DeviceArray((300, 257)) for _ in range(3)]
    stream = hip.Stream()
    DV.boxcar(*[d.ptr for d in din], *[d.ptr for d in dout], 300, 257, 7, stream=stream.handle)
    stream.synchronize()
    for k in range(3):
        assert np.array_equal(dout[k].download(), host[k])
    stream.destroy()
