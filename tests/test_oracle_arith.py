"""Pins the oracle's restatement of the reference's bigfloat add, sub, ucmp,
rcp and div (bigfloat/addsub.c, rcp.c, div.c) with what the mathematics fixes:
exact sums at full precision (rational arithmetic), error bounds of the
truncated sums and of the Newton reciprocal, and known answers derived by hand
from the reference's code: its exponent for short windows, its starting value
built with UINT32_MAX as the word weight, and the sign of a sum of operands
of opposite signs whose second one is the larger in magnitude
(bigfloat/addsub.c:182-190 negates the difference |b| - |a| although it already
carries b's sign, so such a sum comes out with the wrong sign; the reference's
own callers only add operands of one sign). CPU only."""
from fractions import Fraction

import numpy as np
import pytest

from helpers import words_to_int


def value(t):
    w, e, s = t
    v = Fraction(words_to_int(w)) * Fraction(2) ** (32 * e)
    return v if s else -v


def lg2(x):
    """floor-ish log2 of a positive Fraction (exact bit lengths)."""
    return x.numerator.bit_length() - x.denominator.bit_length()


def rand_bf(rng, lo=1, hi=40, exp=20, zero_ok=True):
    n = int(rng.integers(lo, hi + 1))
    if zero_ok and rng.random() < 0.05:
        return (np.zeros(0, np.uint32), 0, 1)
    w = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if rng.random() < 0.2:  # runs of all-ones / zeros: long carries and borrows
        w[: n // 2] = 0xffffffff if rng.random() < 0.5 else 0
    if w[-1] == 0:
        w[-1] = 1
    return (w, int(rng.integers(-exp, exp + 1)), int(rng.integers(0, 2)))


def canonical(t):
    w, e, s = t
    if len(w) == 0:
        assert e == 0 and s == 1, t  # the reference's zero (bigfloat/zero.c)


def reference_sum(a, b):
    """a + b as the reference's bigfloat_add returns it: the exact sum, negated
    when the signs differ and |a| <= |b| (bigfloat/addsub.c:182-190)."""
    x = value(a) + value(b)
    if a[2] != b[2] and abs(value(a)) <= abs(value(b)):
        return -x
    return x


def test_add_sub_exact_at_full_precision(oracle):
    rng = np.random.default_rng(1)
    quirks = 0
    for _ in range(3000):
        a, b = rand_bf(rng, zero_ok=False), rand_bf(rng, zero_ok=False)
        s = oracle.bigfloat_add(a, b, 0)
        d = oracle.bigfloat_sub(a, b, 0)
        assert value(s) == reference_sum(a, b)
        quirks += value(s) != value(a) + value(b)
        assert value(d) == value(a) - value(b)
        canonical(s)
        canonical(d)
    assert quirks > 100


def test_zero_operands(oracle):
    # the reference's zero is (no words, exp 0, sign 1); its magnitude for
    # ucmp is exp + len = 0, as if it were 1 word long at exponent 0
    rng = np.random.default_rng(4)
    zero = (np.zeros(0, np.uint32), 0, 1)
    for _ in range(300):
        b = rand_bf(rng, zero_ok=False)
        b = (b[0], b[1], 1)
        assert value(oracle.bigfloat_add(zero, b, 0)) == value(b)
        assert value(oracle.bigfloat_add(b, zero, 0)) == value(b)
        if b[1] + len(b[0]) > 0:  # |b| >= 1: b beats zero's magnitude in ucmp
            assert value(oracle.bigfloat_sub(b, zero, 0)) == value(b)
    # 0 + (-2^-64): magnitude 0 > -1, so usub(zero, b) over words -2, -1:
    # 0 - 1 borrows through both words: [0xffffffff, 0xffffffff] at exp -2,
    # with zero's sign (positive): 1 - 2^-64
    w, e, s = oracle.bigfloat_add(zero, (np.array([1], np.uint32), -2, 0), 0)
    assert list(map(int, w)) == [0xffffffff, 0xffffffff] and e == -2 and s == 1
    # 0 + (-2^-32): equal magnitudes, b larger: usub(b, zero) takes b's sign and
    # is then negated: +2^-32
    w, e, s = oracle.bigfloat_add(zero, (np.array([1], np.uint32), -1, 0), 0)
    assert list(map(int, w)) == [1] and e == -1 and s == 1


def test_ucmp_is_the_order_of_absolute_values(oracle):
    from oracle.refcpu import RefBigfloat, _u32
    import ctypes
    rng = np.random.default_rng(2)
    L = oracle.lib()
    for _ in range(2000):
        a = rand_bf(rng, 1, 6, 3, zero_ok=False)
        b = (a[0].copy(), a[1], a[2]) if rng.random() < 0.2 else rand_bf(rng, 1, 6, 3, zero_ok=False)
        if rng.random() < 0.3:
            b[0][0] ^= 1
        A, B = RefBigfloat(), RefBigfloat()
        for X, t in ((A, a), (B, b)):
            X.coef, X.exp, X.len, X.sign = _u32(t[0]), t[1], len(t[0]), t[2]
        want = (abs(value(a)) > abs(value(b))) - (abs(value(a)) < abs(value(b)))
        assert L.ref_bigfloat_ucmp(ctypes.byref(A), ctypes.byref(B)) == want


def test_truncated_sums_are_the_window_sums(oracle):
    # with p > 0 the result keeps the top p + 2 words of the magnitude window;
    # each operand loses less than one unit of the window's last word
    rng = np.random.default_rng(3)
    for _ in range(2000):
        a, b = rand_bf(rng, 5, 40, 30, zero_ok=False), rand_bf(rng, 5, 40, 30, zero_ok=False)
        p = int(rng.integers(1, 12))
        top = max(a[1] + len(a[0]), b[1] + len(b[0]))
        bot = min(a[1], b[1])
        if top - bot <= p + 2:
            continue
        unit = Fraction(2) ** (32 * (top - p - 2))
        for op, exact in ((oracle.bigfloat_add, reference_sum(a, b)), (oracle.bigfloat_sub, value(a) - value(b))):
            t = op(a, b, p)
            assert t[1] == top - p - 2 or len(t[0]) == 0
            assert abs(value(t) - exact) < 2 * unit


def test_short_window_exponent_quirk(oracle):
    # bigfloat/addsub.c:78-88,109-110: with p given and the window shorter than
    # p + 2 words, the exponent is still top - (p + 2): 1 + 1 at p = 5 has top 1
    # and one word, so the sum is coef [2] at exponent 1 - 7 = -6
    one = (np.array([1], np.uint32), 0, 1)
    w, e, s = oracle.bigfloat_add(one, one, 5)
    assert list(w) == [2] and e == -6 and s == 1
    # 5 - 3 at p = 1: window of one word, exponent 1 - 3 = -2
    w, e, s = oracle.bigfloat_sub((np.array([5], np.uint32), 0, 1), (np.array([3], np.uint32), 0, 1), 1)
    assert list(w) == [2] and e == -2 and s == 1


def test_sign_rules_and_zero(oracle):
    five, three = (np.array([5], np.uint32), 0, 1), (np.array([3], np.uint32), 0, 1)
    m5 = (five[0], 0, 0)
    assert value(oracle.bigfloat_sub(three, five, 0)) == -2
    assert oracle.bigfloat_sub(three, five, 0)[2] == 0
    assert value(oracle.bigfloat_add(m5, three, 0)) == -2
    # opposite signs, |a| < |b|: -3 + 5 comes out as -2 and 3 + (-5) as +2
    # (the difference 5 - 3 takes b's sign, then bigfloat_neg flips it)
    m3 = (three[0], 0, 0)
    assert value(oracle.bigfloat_add(m3, five, 0)) == -2
    assert value(oracle.bigfloat_add(three, m5, 0)) == 2
    z = oracle.bigfloat_sub(five, five, 0)
    assert len(z[0]) == 0 and z[1] == 0 and z[2] == 1
    z = oracle.bigfloat_add(m5, five, 0)
    assert len(z[0]) == 0 and z[2] == 1  # neg of zero stays positive (bigfloat/neg.c)


def test_rcp_starting_value_known_answers(oracle):
    # bigfloat/rcp.c:30-70 by hand: 1/1 = 1.0 is already >= UINT32_MAX after one
    # scaling by UINT32_MAX: v = 4294967295 -> words (v % UINT32_MAX, v / UINT32_MAX) = (0, 1), exp -1
    assert tuple(map(int, oracle.bigfloat_rcp((np.array([1], np.uint32), 0, 1), 0)[0])) == (0, 1)
    assert oracle.bigfloat_rcp((np.array([1], np.uint32), 0, 1), 0)[1:] == (-1, 1)
    # 1/2: 0.5 * UINT32_MAX = 2147483647.5 < UINT32_MAX, so a second scaling:
    # v = 2147483647.5 * 4294967295 rounded = 2147483647 * 2^32, words
    # (v % UINT32_MAX, v / UINT32_MAX) = (2147483647, 2147483647), exp -2
    w, e, s = oracle.bigfloat_rcp((np.array([2], np.uint32), 0, 0), 0)
    assert list(map(int, w)) == [2147483647, 2147483647] and e == -2 and s == 0


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5, 8, 13, 21, 40, 97])
def test_rcp_error_bound(oracle, p):
    # operands of at least p words (as div and the e pipeline use it):
    # |a r - 1| < 2^-30 for the start value, 2^-(32 p + 28) after Newton
    rng = np.random.default_rng(100 + p)
    bound = Fraction(1, 2 ** 30) if p == 0 else Fraction(1, 2 ** (32 * p + 28))
    for _ in range(60):
        a = rand_bf(rng, max(p, 1), max(p, 1) + 20, 10, zero_ok=False)
        r = oracle.bigfloat_rcp(a, p)
        assert r[2] == a[2]
        assert abs(value(a) * value(r) - 1) < bound, (p, lg2(abs(value(a) * value(r) - 1)))


def test_div_error_bound(oracle):
    rng = np.random.default_rng(7)
    for _ in range(200):
        a, b = rand_bf(rng, 1, 30, 10, zero_ok=False), rand_bf(rng, 1, 30, 10, zero_ok=False)
        t = oracle.bigfloat_div(a, b, 0)
        p = len(a[0]) + len(b[0])
        assert t[2] == (a[2] == b[2])
        q = value(a) / value(b)
        assert abs(value(t) / q - 1) < Fraction(1, 2 ** (32 * min(p, len(b[0])) + 20))
    six, three = (np.array([6], np.uint32), 0, 1), (np.array([3], np.uint32), 0, 1)
    assert abs(value(oracle.bigfloat_div(six, three, 0)) - 2) < Fraction(1, 2 ** 90)
