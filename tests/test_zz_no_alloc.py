"""The allocation-free multiply (bcg_mul_dev_noalloc) through the C ABI
against the allocating one (bcg_mul_dev) on the same operands: two calls in a
row, the second with a smaller input, then the aliased, schoolbook and
one-pass paths, and the size limit."""
import ctypes

import numpy as np
import pytest

from helpers import exact_product
from bluecrunch_amd.workload import operands

pytestmark = pytest.mark.gpu


def _noalloc(bc, c, cl, a, al, b, bl):
    info = bc.MulInfo()
    bc.check(bc.lib().bcg_mul_dev_noalloc(c, cl, a, al, b, bl, None, ctypes.byref(info)), "bcg_mul_dev_noalloc")
    return info.as_dict()


@pytest.mark.parametrize("al,bl", [(300000, 250000), (40000, 1000)])
def test_noalloc_matches_allocating_entry_point(bc, al, bl):
    # first call: three-pass transform near the limit; second: a smaller one
    a, b = operands(al + 3 * bl, al, bl)
    dA, dB = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b)
    dC, dR = bc.DeviceWords(al + bl), bc.DeviceWords(al + bl)
    info = _noalloc(bc, dC.ptr, al + bl, dA.ptr, al, dB.ptr, bl)
    ref = bc.mul_dev(dR.ptr, al + bl, dA.ptr, al, dB.ptr, bl)
    c = dC.to_host()
    assert np.array_equal(c, dR.to_host())
    assert np.array_equal(c, exact_product(a, b))
    assert info["bits"] == ref["bits"] and info["points"] == ref["points"] and info["margin"] < 0.25


def test_noalloc_two_calls_in_a_row_second_smaller(bc):
    big = operands(21, 200000, 200000)
    small = operands(22, 3000, 2000)
    outs = []
    for a, b in (big, small):
        dA, dB = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b)
        dC, dR = bc.DeviceWords(len(a) + len(b)), bc.DeviceWords(len(a) + len(b))
        _noalloc(bc, dC.ptr, len(a) + len(b), dA.ptr, len(a), dB.ptr, len(b))
        bc.mul_dev(dR.ptr, len(a) + len(b), dA.ptr, len(a), dB.ptr, len(b))
        outs.append((dC.to_host(), dR.to_host(), exact_product(a, b)))
    for c, r, x in outs:
        assert np.array_equal(c, r) and np.array_equal(c, x)


@pytest.mark.parametrize("al,bl", [(30, 40), (2000, 3000), (100000, 90000)])
def test_noalloc_aliased_output_and_square(bc, al, bl):
    a, b = operands(al * 7 + bl, al, bl)
    dA = bc.DeviceWords(al + bl)
    bc.check(bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, 4 * al), "h2d")
    dB = bc.DeviceWords.from_host(b)
    _noalloc(bc, dA.ptr, al + bl, dA.ptr, al, dB.ptr, bl)  # C == A
    assert np.array_equal(dA.to_host(), exact_product(a, b))
    dS = bc.DeviceWords(2 * al)
    dA2 = bc.DeviceWords.from_host(a)
    info = _noalloc(bc, dS.ptr, 2 * al, dA2.ptr, al, dA2.ptr, al)
    assert np.array_equal(dS.to_host(), exact_product(a, a))
    assert al + al < 100 or info["square"] == 1


def test_noalloc_rejects_products_over_the_limit(bc):
    w = 1 << 20  # transform of 2^21 or more points
    dA = bc.DeviceWords(w)
    dC = bc.DeviceWords(2 * w)
    info = bc.MulInfo()
    rc = bc.lib().bcg_mul_dev_noalloc(dC.ptr, 2 * w, dA.ptr, w, dA.ptr, w, None, ctypes.byref(info))
    assert rc == -2 and b"limit" in bc.lib().bcg_last_error()
