"""The reference's bigfloat add, sub, mul, rcp and div on the GPU, through the
C ABI (device bigfloats, bcg_bf_*) and through the reference API on host
bigfloats (bigfloat_add/sub/div/rcp), word for word (and exponent, sign)
against the oracle's restatement of bigfloat/addsub.c, mul.c, rcp.c and div.c,
quirks included (short-window exponents, the sign of mixed-sign sums, zero's
magnitude in the comparison)."""
import numpy as np
import pytest

from helpers import words_to_int
from test_oracle_arith import rand_bf

pytestmark = pytest.mark.gpu


def same(x, y):
    return np.array_equal(x[0], y[0]) and x[1] == y[1] and x[2] == y[2]


def dev(bc, name, a, b=None, p=0):
    A = bc.DeviceBigfloat(a)
    B = None if b is None else (A if b is a else bc.DeviceBigfloat(b))
    T = bc.DeviceBigfloat()
    bc.bf_op(name, T, A, B, p)
    return T.triple()


def test_add_sub_random(bc, oracle):
    rng = np.random.default_rng(11)
    for i in range(400):
        a, b = rand_bf(rng), rand_bf(rng)
        p = 0 if i % 3 == 0 else int(rng.integers(1, 45))
        for name in ("add", "sub"):
            want = getattr(oracle, "bigfloat_" + name)(a, b, p)
            got = dev(bc, name, a, b, p)
            assert same(got, want), (name, i, p)
            host = getattr(bc, "bigfloat_" + name)(a, b, p)
            assert same(host, want), (name, i, p)


def test_add_sub_long_carries_and_large(bc, oracle):
    # windows of 10^5 words: the carry resolve of the sums runs over many blocks
    rng = np.random.default_rng(12)
    n = 100000
    ones = np.full(n, 0xffffffff, np.uint32)
    one = np.array([1], np.uint32)
    cases = [((ones, 0, 1), (one, 0, 1)), ((ones, -5, 1), (one, -5, 0)), ((one, n, 1), (ones, 0, 0))]
    for _ in range(3):
        cases.append((rand_bf(rng, n - 100, n, 50, zero_ok=False), rand_bf(rng, n - 100, n, 50, zero_ok=False)))
    for a, b in cases:
        for name in ("add", "sub"):
            for p in (0, 1000, n // 2):
                want = getattr(oracle, "bigfloat_" + name)(a, b, p)
                assert same(dev(bc, name, a, b, p), want), (name, p)


def test_add_sub_of_nearly_equal_magnitudes(bc, oracle):
    # |a| and |b| agree on every word above one position: deep in a 10^5-word
    # window, at its top, nowhere, or only in b's extra low word (b one word
    # longer, one exponent lower). The magnitude comparison of a same-sign sub
    # and an opposite-sign add scans down to that position.
    rng = np.random.default_rng(14)
    n = 100000
    w = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    w[-1] |= np.uint32(1)
    cases = []
    for k in (0, 3, 300, 70000, n - 1, None):
        v = w.copy()
        if k is not None:
            v[k] ^= np.uint32(2)  # the top word keeps its bit 0
        cases.append(((w, 3, 1), (v, 3, 1)))
    cases.append(((w, 3, 1), (np.concatenate([np.array([5], np.uint32), w]), 2, 1)))
    for a, b in cases:
        for name, bb in (("sub", b), ("add", (b[0], b[1], 1 - b[2]))):
            for p in (0, 1000):
                want = getattr(oracle, "bigfloat_" + name)(a, bb, p)
                assert same(dev(bc, name, a, bb, p), want), (name, p)


def test_in_place_targets(bc, oracle):
    rng = np.random.default_rng(13)
    for i in range(100):
        a, b = rand_bf(rng, zero_ok=False), rand_bf(rng, zero_ok=False)
        b = (b[0], a[1], b[2])  # windows starting at a's exponent: in-place is well defined in the reference
        for name in ("add", "sub", "mul"):
            want = getattr(oracle, "bigfloat_" + name)(a, b, 0)
            A, B = bc.DeviceBigfloat(a), bc.DeviceBigfloat(b)
            bc.bf_op(name, A, A, B, 0)  # t = a
            assert same(A.triple(), want), (name, i)
            assert same(getattr(bc, "bigfloat_" + name)(a, b, 0, inplace="a"), want), (name, i)


@pytest.mark.parametrize("p", [0, 1, 3, 40, 2000])
def test_mul_truncation(bc, oracle, p):
    rng = np.random.default_rng(14 + p)
    for _ in range(20):
        a, b = rand_bf(rng, 1, 3000, 30), rand_bf(rng, 1, 3000, 30)
        want = oracle.bigfloat_mul(a, b, p)
        assert same(dev(bc, "mul", a, b, p), want)


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5, 8, 13, 40, 97, 300, 1000])
def test_rcp_matches_oracle(bc, oracle, p):
    rng = np.random.default_rng(15 + p)
    for k in range(12):
        # operands shorter and longer than p (the short-window exponent applies to the first)
        lo, hi = (1, 4) if k < 4 else (max(p, 1), max(p, 1) + 30)
        a = rand_bf(rng, lo, hi, 10, zero_ok=False)
        want = oracle.bigfloat_rcp(a, p)
        assert same(dev(bc, "rcp", a, None, p), want), (p, k)
        assert same(bc.bigfloat_rcp(a, p), want), (p, k)


@pytest.mark.parametrize("p", [20000, 120000, 400000])
def test_rcp_large_matches_oracle(bc, oracle, p):
    # precisions whose Newton products run as multi-pass transforms
    rng = np.random.default_rng(17 + p)
    for k in range(2):
        a = rand_bf(rng, p, p + 30, 10, zero_ok=False) if k else rand_bf(rng, p // 3, p // 2, 10, zero_ok=False)
        want = oracle.bigfloat_rcp(a, p)
        assert same(dev(bc, "rcp", a, None, p), want), (p, k)


@pytest.mark.parametrize("words", [50, 700, 5000, 60000, 250000])
def test_div_matches_oracle(bc, oracle, words):
    rng = np.random.default_rng(16 + words)
    for _ in range(3):
        a = rand_bf(rng, words // 2, words, 10, zero_ok=False)
        b = rand_bf(rng, words // 2, words, 10, zero_ok=False)
        for p in (0, words // 3 + 1):
            want = oracle.bigfloat_div(a, b, p)
            assert same(dev(bc, "div", a, b, p), want), p
            assert same(bc.bigfloat_div(a, b, p), want), p


def test_e_division_step(bc, oracle):
    # main.c:131-141 at 10^5 digits: e - 1 = P/Q to p words, then + 1
    from bluecrunch_amd.workload import words_to_int as wi
    P, Q = bc.e_bsr(bc.e_terms(11113))
    p = 11113
    a, b = (P, 0, 1), (Q, 0, 1)
    want = oracle.bigfloat_div(a, b, p)
    got = dev(bc, "div", a, b, p)
    assert same(got, want)
    one = (np.array([1], np.uint32), 0, 1)
    e_host = bc.bigfloat_add(got, one, p)
    assert same(e_host, oracle.bigfloat_add(want, one, p))
    # the leading words are e's: 2.718281828... as a fraction of 2^32
    assert e_host[0][-1] == 2 and int(e_host[0][-2]) == int(0.718281828459045 * 2**32)


def test_zero_and_argument_errors(bc):
    zero = (np.zeros(0, np.uint32), 0, 1)
    five = (np.array([5], np.uint32), 0, 1)
    assert dev(bc, "mul", zero, five)[0].size == 0
    with pytest.raises(bc.GpuError):
        dev(bc, "rcp", zero, None, 3)
