"""fr_sum_noalloc (the allocation-free variant of fr_sum, its scratch in a static device array)
through the C-ABI against fr_sum: the same result bits on two calls in a row, the second on a
smaller input, and its argument checks."""
import ctypes as C
import sys

import numpy as np
import pytest

import fastreg_amd as fr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sums(L, ptr, n):
    a, b = C.c_double(), C.c_double()
    fr.check(L.fr_sum(ptr, n, C.byref(a), None))
    fr.check(L.fr_sum_noalloc(ptr, n, C.byref(b), None))
    return a.value, b.value


def test_sum_noalloc_equals_sum_on_two_calls():
    L = fr.lib()
    rng = np.random.default_rng(71)
    # the first call uses all 1024 partial slots (n >= 256 * 1023); the second, smaller one 40
    for n in (3_000_001, 10_000):
        say("fr_sum_noalloc", n)
        x = (rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, n)).astype(np.float32)
        d = fr.DeviceArray.from_host(x)
        try:
            ref, got = _sums(L, d.ptr, n)
        finally:
            d.free()
        assert np.float64(got).view(np.uint64) == np.float64(ref).view(np.uint64), (n, got, ref)
        np.testing.assert_allclose(got, np.sum(x.astype(np.float64)), rtol=1e-12)


def test_sum_noalloc_small_and_empty_inputs():
    L = fr.lib()
    x = np.arange(1, 8, dtype=np.float32)
    d = fr.DeviceArray.from_host(x)
    try:
        out = C.c_double()
        fr.check(L.fr_sum_noalloc(d.ptr, 7, C.byref(out), None))
        assert out.value == 28.0
        fr.check(L.fr_sum_noalloc(d.ptr, 0, C.byref(out), None))
        assert out.value == 0.0
        assert L.fr_sum_noalloc(d.ptr, -1, C.byref(out), None) == fr.FR_ERR_INVALID
        assert L.fr_sum_noalloc(None, 7, C.byref(out), None) == fr.FR_ERR_INVALID
    finally:
        d.free()
