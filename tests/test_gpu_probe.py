"""The read-bandwidth probe used for the cache rooflines: it runs, times itself and checks its
sink size; a cache-resident buffer reads faster than one that must come from HBM."""
import ctypes as C

import pytest

from imbatracer_amd import native as N

pytestmark = pytest.mark.gpu


def _gbs(size, passes, sink):
    buf = N.DeviceBuffer(size)
    N.check(N.lib.imba_memset(buf.ptr, 0, size, None), "memset")
    ms = C.c_float()
    best = 0.0
    for _ in range(3):
        N.check(N.lib.imba_read_bandwidth(buf.ptr, size, passes, sink.ptr, sink.nbytes, None, C.byref(ms)), "probe")
        assert ms.value > 0
        best = max(best, size * passes / (ms.value * 1e-3) / 1e9)
    return best


def test_read_bandwidth(gpu):
    sink = N.DeviceBuffer(32 * 1024 * 256)
    l2 = _gbs(2 << 20, 500, sink)
    hbm = _gbs(1 << 30, 2, sink)
    assert 500 < hbm < 9000  # GB/s: below the 8 TB/s HBM peak, well above PCIe
    assert l2 > hbm
    small = N.DeviceBuffer(64)
    buf = N.DeviceBuffer(1 << 20)
    ms = C.c_float()
    assert N.lib.imba_read_bandwidth(buf.ptr, 1 << 20, 1, small.ptr, small.nbytes, None, C.byref(ms)) == 5  # capacity
    assert N.lib.imba_read_bandwidth(buf.ptr, 1 << 20, 0, sink.ptr, sink.nbytes, None, C.byref(ms)) == 1  # passes
