"""Radix sort and scan on the GPU against numpy (stable argsort, cumsum),
as test/unit/algorithm/radix_sort.cpp and test/unit/algorithm/gpu/scan.cpp do
against std::sort and a serial prefix sum.  Bit-exact."""
import ctypes

import numpy as np
import pytest

from halmd_amd import _lib

pytestmark = pytest.mark.gpu


class DeviceArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        _lib.call("halmd_device_alloc", max(self.host.nbytes, 4), ctypes.byref(self.ptr))
        if self.host.nbytes:
            _lib.call("halmd_memcpy_htod", self.ptr, self.host.ctypes.data_as(ctypes.c_void_p), self.host.nbytes)

    def get(self):
        out = np.empty_like(self.host)
        if out.nbytes:
            _lib.call("halmd_memcpy_dtoh", out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes)
        return out

    def __del__(self):
        if self.ptr.value:
            _lib.call("halmd_device_free", self.ptr)


@pytest.mark.parametrize("n,bits", [(0, 8), (1, 8), (7, 3), (4095, 16), (4096, 16), (4097, 12), (100003, 21),
                                    (1048576, 17), (3000017, 32)])
def test_radix_sort_is_stable_sort(n, bits):
    rng = np.random.default_rng(n + bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[: n // 3] = keys[0]  # many equal keys: stability matters
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = DeviceArray(keys), DeviceArray(vals)
    _lib.call("halmd_algorithm_radix_sort", dk.ptr, dv.ptr, n, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.get(), keys[order])
    assert np.array_equal(dv.get(), vals[order])


@pytest.mark.parametrize("n", [1, 255, 2048, 2049, 100000, 4194305, 10_000_000])
def test_scan_exclusive(n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 300, n, dtype=np.uint32)
    d = DeviceArray(x)
    _lib.call("halmd_algorithm_scan_exclusive", d.ptr, n)
    ref = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(d.get(), ref)
