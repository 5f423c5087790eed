"""The kernels' reciprocal (hardware approximation refined by one FMA Newton
step) equals the correctly rounded division 1.0f/x for every positive FP32
x in [2^-125, 2^125) -- checked exhaustively, bit pattern by bit pattern, on
the device.  The SOR's cn1 = 1/((ax+by)+cz) relies on it (init.f95:100-107
values lie far inside that range)."""
import ctypes

import pytest

from lesamd import api

pytestmark = pytest.mark.gpu


def test_reciprocal_is_correctly_rounded_over_the_normal_range():
    L = api.lib()
    L.les_selftest_reciprocal.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]
    lo, hi = 2 << 23, 252 << 23          # bit patterns of 2^-125 and 2^125
    miss, first = ctypes.c_ulonglong(0), ctypes.c_uint(0)
    print(f"\n[reciprocal self-test] {hi - lo} values", flush=True)
    assert L.les_selftest_reciprocal(lo, hi, ctypes.byref(miss), ctypes.byref(first)) == 0, \
        L.les_last_error().decode()
    assert miss.value == 0, (miss.value, hex(first.value))
