"""Deferred margin checks (bcg_defer_begin / bcg_defer_end): products queued
back to back without a read-back per product stay exact, the scope reports
their worst margin, a scope whose margins reach 0.25 is detected, and the
calls outside any scope are unchanged."""
import numpy as np
import pytest

from helpers import exact_product
from bluecrunch_amd.workload import operands

pytestmark = pytest.mark.gpu


def _queue(bc, pairs):
    bufs = []
    for a, b in pairs:
        dA, dB = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b)
        dC = bc.DeviceWords(len(a) + len(b))
        info = bc.mul_dev(dC.ptr, len(a) + len(b), dA.ptr, len(a), dB.ptr, len(b))
        bufs.append((dA, dB, dC, info))
    return bufs


def test_scope_of_products_is_exact(bc):
    # one-pass, two- and three-pass lengths, more than one set of margin slots
    sizes = [(300, 200), (5000, 4000), (40000, 30000), (200000, 150000)] * 20
    pairs = [operands(11 + k, al, bl) for k, (al, bl) in enumerate(sizes)]
    bc.stats_reset()
    bc.defer_begin()
    bufs = _queue(bc, pairs)
    worst = bc.defer_end()
    assert 0.0 < worst < 0.25
    for (a, b), (_, _, dC, info) in zip(pairs, bufs):
        assert info["margin"] == 0.0  # not read inside the scope
        assert np.array_equal(dC.to_host(), exact_product(a, b))
    st = bc.stats()
    assert st["products"] == len(pairs) and abs(st["worst_margin"] - worst) < 1e-12


def test_scope_detects_a_margin_over_the_bound(bc):
    # digit sizes planned 6 bits over the bound: the scope must report it,
    # and the same products outside a scope re-run themselves and are exact
    pairs = [operands(31 + k, 60000, 50000) for k in range(3)]
    bc.plan_digit_bias(6)
    try:
        bc.defer_begin()
        _queue(bc, pairs)
        worst = bc.defer_end()
        assert worst >= 0.25
        for a, b in pairs:
            c, info = bc.mul(a, b)
            assert info["retries"] >= 1 and np.array_equal(c, exact_product(a, b))
    finally:
        bc.plan_digit_bias(0)


def test_aliased_product_in_a_scope_checks_at_once(bc):
    a, b = operands(41, 30000, 30000)
    dA, dB = bc.DeviceWords(60000), bc.DeviceWords.from_host(b)
    bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, a.nbytes)
    bc.defer_begin()
    try:
        info = bc.mul_dev(dA.ptr, 60000, dA.ptr, 30000, dB.ptr, 30000)
    finally:
        bc.defer_end()
    assert info["margin"] > 0.0  # read at once: the retry needs the inputs
    assert np.array_equal(dA.to_host(), exact_product(a, b))


def test_scope_misuse_is_an_error(bc):
    with pytest.raises(bc.GpuError):
        bc.defer_end()
    bc.defer_begin()
    try:
        with pytest.raises(bc.GpuError):
            bc.defer_begin()
    finally:
        bc.defer_end()
