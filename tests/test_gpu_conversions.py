"""The device float -> int conversions every kernel uses (truncation toward
zero, saturating, NaN -> 0) are single hardware instructions; their equality
with the definition is checked on the device over all 2^32 float inputs, and
a strided sample of the device results is checked here against numpy's
truncation with the same saturation and NaN rule."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_conversions_match_definition_on_every_input(device):
    from apbf_amd import _lib
    L = _lib.lib()
    stride = 4099
    samples = ((1 << 32) - 1) // stride + 1
    si = np.zeros(samples, np.int32)
    su = np.zeros(samples, np.uint32)
    bad = C.c_uint64(123)
    first = C.c_uint32(0)
    _lib.check(L.apbf_check_conversions(stride, C.byref(bad), C.byref(first), si.ctypes.data_as(C.c_void_p),
                                        su.ctypes.data_as(C.c_void_p)), "apbf_check_conversions")
    assert bad.value == 0, f"{bad.value} patterns differ, first 0x{first.value:08x}"

    bits = (np.arange(samples, dtype=np.uint64) * stride).astype(np.uint32)
    with np.errstate(invalid="ignore"):  # signalling NaNs among the patterns
        x = bits.view(np.float32).astype(np.float64)
    t = np.trunc(np.nan_to_num(x, nan=0.0, posinf=np.inf, neginf=-np.inf))
    want_i = np.clip(t, -2.0 ** 31, 2.0 ** 31 - 1).astype(np.int64)
    want_u = np.clip(t, 0.0, 2.0 ** 32 - 1).astype(np.int64)
    assert np.array_equal(si.astype(np.int64), want_i)
    assert np.array_equal(su.astype(np.int64), want_u)
    # the sample holds NaNs, values beyond both saturation limits and values below 1 in magnitude
    assert np.isnan(x).any() and (x > 2.0 ** 32).any() and (x < -2.0 ** 31).any() and (np.abs(x) < 1).any()


def test_short_sqrt_and_reciprocal_equal_ieee(device):
    """sqrt_in_range and recip_in_range (hardware approximation plus one
    correction) give the IEEE bits on every float of their argument ranges,
    and quotient_by_recip (the solver's division by inverse masses) gives the
    IEEE quotient for 512 divisors of 2^-40 to 2^40, each with every dividend
    significand."""
    from apbf_amd import _lib
    L = _lib.lib()
    bad = np.full(3, 7, np.uint64)
    first = np.zeros(3, np.uint32)
    _lib.check(L.apbf_check_math(bad.ctypes.data_as(C.c_void_p), first.ctypes.data_as(C.c_void_p)), "apbf_check_math")
    names = ("quotient", "sqrt", "reciprocal")
    assert not bad.any(), {names[f]: (int(bad[f]), hex(first[f])) for f in range(3) if bad[f]}
