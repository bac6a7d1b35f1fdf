"""The radiative-loss branch bounds (radiativelosses.cpp:158-180 selects the loss-curve branch by
comparing log10(T), the C library's, with 4.97, 5.67, 6.18, 6.55, 6.90 and 7.63).  The GPU compares
T with the largest double whose host log10 is at most each constant, formed at set-up the same way
as here; that selects the reference's branch for every T exactly when the host log10 is monotone
across the bound, which is checked here over 2^16 doubles on either side of each."""
import math
import struct

LIM = [4.97, 5.67, 6.18, 6.55, 6.90, 7.63]


def upper(c):
    t = math.pow(10.0, c)
    while math.log10(t) > c:
        t = math.nextafter(t, 0.0)
    while math.log10(math.nextafter(t, math.inf)) <= c:
        t = math.nextafter(t, math.inf)
    return t


def _step(x, k):
    """The double k representable steps above (k > 0) or below x (positive x)."""
    b = struct.unpack("<q", struct.pack("<d", x))[0] + k
    return struct.unpack("<d", struct.pack("<q", b))[0]


def test_host_log10_is_monotone_across_every_bound():
    for c in LIM:
        tb = upper(c)
        assert math.log10(tb) <= c < math.log10(math.nextafter(tb, math.inf))
        assert abs(tb / 10.0 ** c - 1.0) < 1e-14
        for k in range(-(1 << 16), (1 << 16) + 1):
            x = _step(tb, k)
            assert (math.log10(x) <= c) == (x <= tb), (c, k)
