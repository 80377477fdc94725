"""Binary splitting of e on the GPU against exact integers
(docs/binary-splitting.txt example, closed forms P = sum (k+1)...T, Q = T!)."""
import math

import numpy as np
import pytest

from helpers import words_to_int

pytestmark = pytest.mark.gpu


def exact_pq(T):
    q = math.factorial(T)
    p, run = 0, 1
    for k in range(T, 0, -1):  # P = sum_{k=1}^{T} (k+1)...T
        p += run
        run *= k
    return p, q


def test_bsr_docs_example(bc):
    # docs/binary-splitting.txt:44-92: 5 terms -> P = 206, Q = 120
    P, Q = bc.e_bsr(5)
    assert words_to_int(P) == 206 and words_to_int(Q) == 120


@pytest.mark.parametrize("T", [1, 2, 15, 16, 17, 33, 100, 1000, 4097, 30000, 200000])
def test_bsr_exact(bc, T):
    P, Q = bc.e_bsr(T)
    p, q = exact_pq(T)
    assert words_to_int(Q) == q
    assert words_to_int(P) == p
    assert P[-1] != 0 and Q[-1] != 0


@pytest.mark.parametrize("T,nw", [(5, 2), (50, 4), (1000, 40), (30000, 900), (200000, 6000)])
def test_fraction_division(bc, T, nw):
    # fractional part of P/Q (= e - 2 for large T): divide P - Q by Q
    from helpers import int_to_words
    P, Q = bc.e_bsr(T)
    p, q = words_to_int(P), words_to_int(Q)
    pm = p - q
    assert 0 <= pm < q
    n = max(1, (pm.bit_length() + 31) // 32)
    F = bc.div_fraction(int_to_words(pm, n), Q, nw)
    want = (pm << (32 * nw)) // q
    assert words_to_int(F) == want


def e_digits_exact(D):
    """Truncated decimal digits of e by integer arithmetic: floor(e 10^D)."""
    import sys
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(0)
    scale = 10 ** (D + 10)
    s, term, k = 0, scale, 0
    while term:
        s += term
        k += 1
        term //= k
    return str(s // 10 ** 10)


@pytest.mark.parametrize("groups", [1, 5, 31, 32, 33, 100, 1000, 5000])
def test_radix_exact(bc, groups):
    # fraction 1/7-ish pattern and a random one, against Python's exact digits
    rng = np.random.default_rng(groups)
    n = int(np.ceil(groups * 0.9343253125117121)) + 3
    F = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    got = bc.radix(F, groups)
    x = words_to_int(F)
    want = []
    for _ in range(groups):
        x *= 10 ** 9
        want.append(x >> (32 * n))
        x &= (1 << (32 * n)) - 1
    assert list(map(int, got)) == want


@pytest.mark.parametrize("D", [1000, 10000, 100000])
def test_e_digits_md5(bc, golden, D):
    import hashlib
    txt, ms = bc.e_digits(D)
    assert len(txt) == D + 2
    assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"][str(D)]
    if D <= 10000:
        assert txt == "2." + e_digits_exact(D)[1:]


@pytest.mark.gpu
@pytest.mark.parametrize("D", [31337, 123457])
def test_e_digits_irregular_counts_exact(bc, D):
    # digit counts that are not powers of ten (a partial last group of nine
    # digits, odd tree shapes) against floor(e 10^D) from integer arithmetic
    txt, ms = bc.e_digits(D)
    assert txt == "2." + e_digits_exact(D)[1:]


def test_e_digits_1e6_md5(bc, golden):
    import hashlib
    txt, ms = bc.e_digits(1000000)
    assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"]["1000000"]


@pytest.mark.gpu
@pytest.mark.parametrize("D", [1000, 100000])
def test_e_digits_per_multiply_path(bc, golden, D):
    # the plain per-multiply binary splitting gives the same digits as the batched one
    import hashlib
    txt, ms = bc.e_digits(D, per_multiply=True)
    assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"][str(D)]


@pytest.mark.gpu
def test_e_digits_1e7_md5(bc, golden):
    # GPU-side decimal formatting at 1e7 digits (raw bytes, reference md5sums)
    import hashlib
    txt, ms = bc.e_digits(10 ** 7, raw=True)
    assert len(txt) == 10 ** 7 + 2 and txt[:2] == b"2."
    assert hashlib.md5(txt).hexdigest() == golden["e_md5"]["10000000"]


@pytest.mark.gpu
def test_e_digits_1e9_md5(bc, golden):
    # the largest configuration: 1e9 digits (products up to ~2^29 complex points)
    import hashlib
    txt, ms = bc.e_digits(10 ** 9, raw=True)
    assert hashlib.md5(txt).hexdigest() == golden["e_md5"]["1000000000"]


def test_error_paths_leave_the_pool_consistent(bc, golden):
    # every device allocation of the e pipeline failed in turn (binary
    # splitting levels, division, radix conversion): the call reports the
    # error, and the next call through the same memory pool is still exact
    import ctypes
    import hashlib
    L = bc.lib()
    buf = np.empty(10 ** 4 + 3, dtype=np.uint8)
    ms = (ctypes.c_double * 4)()
    failed = 0
    for nth in range(1, 400):
        bc.check(L.bcg_test_fail_alloc(nth), "fail_alloc")
        rc = L.bcg_e_digits_ex(10 ** 4, buf.ctypes.data_as(ctypes.c_char_p), buf.size, ms, 0)
        bc.check(L.bcg_test_fail_alloc(0), "fail_alloc")
        if rc == 0:
            break
        failed += 1
        assert rc < 0
        txt, _ = bc.e_digits(10 ** 4)
        assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"]["10000"], nth
    assert failed >= 10 and rc == 0
    # radix conversion alone, on a random fraction
    rng = np.random.default_rng(5)
    F = rng.integers(0, 2**32, 4000, dtype=np.uint64).astype(np.uint32)
    want = bc.radix(F, 4000)
    dF, dD = bc.DeviceWords.from_host(F), bc.DeviceWords(4000)
    for nth in range(1, 400):
        bc.check(L.bcg_test_fail_alloc(nth), "fail_alloc")
        rc = L.bcg_radix_dev(dF.ptr, 4000, 4000, dD.ptr)
        bc.check(L.bcg_test_fail_alloc(0), "fail_alloc")
        if rc == 0:
            break
        assert np.array_equal(bc.radix(F, 4000), want), nth
    assert rc == 0 and np.array_equal(dD.to_host(), want)


def test_e_text_equals_the_oracle_for_small_digit_counts(bc, oracle):
    # e.txt as the reference writes it, including its short texts
    for D in list(range(1, 60)) + [99, 100, 101, 997, 2000]:
        txt, _ = bc.e_digits(D)
        assert txt.encode() == oracle.e_digits(D)[0], D
