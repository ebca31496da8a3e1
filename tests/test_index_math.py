"""Index arithmetic of the NL-SAR scale pass, checked exhaustively on the host.

The pass forms every region row / run index as q = trunc((n + 1/2) * rcp(d)) in float32,
with rcp(d) from v_rcp_f32, which is within 1 ulp of 1/d.  The claim (exact for
0 <= n < 2^16 and 1 <= d <= 1024, the ranges of the pass's regions) is checked here for the
correctly rounded reciprocal and for its neighbours one ulp above and below, which covers
every value the hardware reciprocal can return.  A mistaken form (n instead of n + 1/2, or
a plain float division of n) fails it.
"""
import numpy as np

N = np.arange(1 << 16, dtype=np.int64)


def _quotients(d, rcp, half=0.5):
    num = (N.astype(np.float32) + np.float32(half)).astype(np.float32)
    return np.floor((num * rcp).astype(np.float32)).astype(np.int64)


def _reciprocals(d):
    r = np.float32(1.0) / np.float32(d)
    return (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(np.inf)))


def test_index_quotient_exact_for_every_hardware_reciprocal():
    for d in range(1, 1025):
        exact = N // d
        for rcp in _reciprocals(d):
            q = _quotients(d, rcp)
            assert np.array_equal(q, exact), f"d {d}, rcp {rcp!r}"


def test_index_quotient_needs_the_half_offset():
    """Without the 1/2 the product of n = k*d with a reciprocal below 1/d truncates to k - 1."""
    bad = 0
    for d in (3, 7, 11, 38, 43, 48, 62):
        rcp = _reciprocals(d)[0]
        bad += int((_quotients(d, rcp, half=0.0) != N // d).sum())
    assert bad > 0
