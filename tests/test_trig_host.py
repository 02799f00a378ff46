"""The sin/cos table path on the host, against the host libm for every float of the domain.

The kernels evaluate sin/cos in double precision, round to f32 and search the exception list
only when the double value lies within the host-measured distance of a rounding midpoint; the host evaluation here runs the
same code (dl_trig.h), so an exhaustive pass pins the lookup logic without a GPU.  The device
side is compared with this host path in tests/test_gpu_parity.py.
"""
import ctypes as C

import pytest

import dynlight_hip


@pytest.fixture(scope="module")
def lib():
    try:
        return dynlight_hip.load()
    except OSError as e:
        pytest.skip(str(e))


@pytest.mark.parametrize("lo,hi", [(0x37800000, 0x41000000), (0xB7800000, 0xBF4CCCCD)])
def test_table_path_equals_libm_on_the_whole_domain(lib, lo, hi):
    bad = C.c_uint64()
    assert lib.dl_test_trig_exhaustive(lo, hi, C.byref(bad)) == 0
    assert bad.value == 0


def test_exceptions_exist_and_lie_in_the_domain(lib):
    for which in (0, 1):
        n = lib.dl_test_trig_keys(which, None, 0)
        keys = (C.c_uint32 * n)()
        lib.dl_test_trig_keys(which, keys, n)
        # glibc rounds about 1% of sin and 0.2% of cos results differently from the nearest float
        assert n > 100_000
        assert all((0x37800000 <= k <= 0x41000000) or (0xB7800000 <= k <= 0xBF4CCCCD) for k in keys[:: max(1, n // 5000)])
        assert list(keys[:1000]) == sorted(keys[:1000])
