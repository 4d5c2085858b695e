"""The allocation-free radix sort (halmd_algorithm_radix_sort_noalloc) against
the allocating entry point on the same inputs: bit-exact, on two calls in a
row (the second smaller, on the scratch the first left), and the size limit."""
import ctypes

import numpy as np
import pytest

from halmd_amd import _lib
from test_gpu_primitives import DeviceArray

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 21


def _sorted_by(entry, keys, vals, bits):
    dk, dv = DeviceArray(keys), DeviceArray(vals)
    _lib.call(entry, dk.ptr, dv.ptr, len(keys), bits)
    return dk.get(), dv.get()


@pytest.mark.parametrize("sizes,bits", [((NOALLOC_MAX, 300001), 21), ((1048576, 4097), 17), ((70001, 5), 32)])
def test_noalloc_sort_equals_allocating_sort(sizes, bits):
    rng = np.random.default_rng(bits)
    for n in sizes:
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 4] = keys[n // 2]     # long runs of equal keys: stability shows
        vals = rng.permutation(n).astype(np.uint32)
        k0, v0 = _sorted_by("halmd_algorithm_radix_sort", keys, vals, bits)
        k1, v1 = _sorted_by("halmd_algorithm_radix_sort_noalloc", keys, vals, bits)
        assert np.array_equal(k1, k0) and np.array_equal(v1, v0)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(v1, vals[order])


def test_noalloc_sort_rejects_more_than_its_scratch():
    # rejected on the host before any launch (the pointers are never touched)
    dk, dv = DeviceArray(np.zeros(4, np.uint32)), DeviceArray(np.zeros(4, np.uint32))
    rc = _lib.lib().halmd_algorithm_radix_sort_noalloc(dk.ptr, dv.ptr, ctypes.c_size_t(NOALLOC_MAX + 1),
                                                      ctypes.c_uint(8))
    assert rc == 3   # HALMD_ERR_CAPACITY
    assert b"scratch" in _lib.lib().halmd_last_error()
