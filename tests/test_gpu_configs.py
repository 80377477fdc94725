"""The benchmark configurations as parity tests on the GPU, through the C ABI:

* the plain multiply at 2^17 words with the reference's own digit size (b = 9),
  word for word against the oracle;
* the 2^24-word product and square, word for word against the oracle;
* the 4096 uneven products (seed 7) captured in one hipGraph and replayed,
  every product exact against Python integers;
* e to 10^8 (and 2*10^8) digits against the reference's md5sums;
* the rounding-margin safety net: products planned with too many bits per digit
  must re-run with smaller digits (retries counted) and stay exact, for the
  single one-pass and multi-pass paths, aliased and not, batches and
  multiply-adds, and the whole e pipeline.
"""
import ctypes
import hashlib

import numpy as np
import pytest

from helpers import exact_product, words_to_int
from bluecrunch_amd.workload import batch_operands, operands

pytestmark = pytest.mark.gpu


@pytest.fixture
def no_bias(bc):
    yield
    bc.plan_digit_bias(0)


def test_config1_reference_digit_size(bc, oracle):
    # 2^17 x 2^17 words at the reference's schedule for CL = 2^18 words
    w = 1 << 17
    a, b = operands(1, w, w)
    bits = oracle.lib().ref_sched_bits(2 * w)
    assert bits == 9
    dA, dB, dC = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b), bc.DeviceWords(2 * w)
    info = bc.mul_dev(dC.ptr, 2 * w, dA.ptr, w, dB.ptr, w, bits=bits)
    assert info["bits"] == 9 and info["retries"] == 0 and info["margin"] < 0.25
    want, obits, _ = oracle.fft_mul(a, b)
    assert obits == 9
    assert np.array_equal(dC.to_host(), want)


def test_config2_product_and_square_word_for_word(bc, oracle):
    w = 1 << 24
    a, b = operands(1, w, w)
    dA, dB = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b)
    dC = bc.DeviceWords(2 * w)
    info = bc.mul_dev(dC.ptr, 2 * w, dA.ptr, w, dB.ptr, w)
    assert info["bits"] == 8 and info["points"] == 1 << 26 and info["margin"] < 0.25
    c = dC.to_host()
    want, _, _ = oracle.fft_mul(a, b)
    assert np.array_equal(c, want)
    del c, want
    sinfo = bc.mul_dev(dC.ptr, 2 * w, dA.ptr, w, dA.ptr, w)
    assert sinfo["square"] == 1 and sinfo["margin"] < 0.25
    s = dC.to_host()
    want, _, _ = oracle.fft_mul(a, a)
    assert np.array_equal(s, want)


def test_config3_batch_graph_4096_exact(bc):
    pairs = batch_operands(7, 4096)
    al = np.array([len(x) for x, _ in pairs], dtype=np.uint64)
    bl = np.array([len(y) for _, y in pairs], dtype=np.uint64)
    cl = al + bl
    off = lambda n: np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)  # noqa: E731
    a_off, b_off, c_off = off(al), off(bl), off(cl)
    dA = bc.DeviceWords.from_host(np.concatenate([x for x, _ in pairs]))
    dB = bc.DeviceWords.from_host(np.concatenate([y for _, y in pairs]))
    dC = bc.DeviceWords(int(cl.sum()))
    want = [exact_product(x, y) for x, y in pairs]
    batch = bc.Batch(dC, dA, dB, c_off, cl, a_off, al, b_off, bl)
    try:
        for rep in range(2):
            dC.zero()
            worst = batch.run()
            C = dC.to_host()
            bad = [i for i, w in enumerate(want) if not np.array_equal(C[int(c_off[i]):int(c_off[i] + cl[i])], w)]
            assert not bad, (rep, bad[:10])
            assert worst < 0.25
    finally:
        batch.close()


@pytest.mark.parametrize("digits", [10 ** 8, 2 * 10 ** 8])
def test_e_digits_md5_large(bc, golden, digits):
    txt, ms = bc.e_digits(digits, raw=True)
    assert len(txt) == digits + 2 and txt[:2] == b"2."
    assert hashlib.md5(txt).hexdigest() == golden["e_md5"][str(digits)]


# ------------------------------------------------------- margin safety net

@pytest.mark.parametrize("words", [5000, 50000])
@pytest.mark.parametrize("alias", [False, True])
def test_margin_retry_single(bc, words, alias):
    # 20 bits per digit, far above the bound the planner allows at these
    # lengths (one-pass N = 8192 at 5000 words, three-pass N = 3*2^15 at
    # 50000): FP64 errors of several units in the large coefficients and
    # random fractions in the small ones at the ends of the convolution, so the
    # first attempt reaches a margin >= 0.25 and re-runs follow with fewer bits
    a, b = operands(300 + words, words, words)
    assert bc.plan(words, words)["bits"] <= 16
    dA = bc.DeviceWords(2 * words)
    bc.check(bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, 4 * words), "h2d")
    dB = bc.DeviceWords.from_host(b)
    dC = dA if alias else bc.DeviceWords(2 * words)
    bc.stats_reset()
    info = bc.mul_dev(dC.ptr, 2 * words, dA.ptr, words, dB.ptr, words, bits=20)
    assert info["retries"] >= 1 and info["bits"] < 20 and info["margin"] < 0.25, info
    assert bc.stats()["retries"] == info["retries"]
    assert np.array_equal(dC.to_host(), exact_product(a, b))


def test_margin_retry_batch_and_muladd(bc, no_bias):
    from bluecrunch_amd.workload import SplitMix64, operand
    rng = SplitMix64(77)
    # one-pass classes and multi-pass groups; with 6 bits of bias the larger
    # ones are planned at (up to) 20 bits per digit and must re-run
    sizes = [(300, 200), (2000, 3000), (5000, 5000), (7000, 7000), (20000, 9000), (40000, 40000)]
    pairs = [(operand(rng, x), operand(rng, y)) for x, y in sizes]
    adds = [operand(rng, 50) for _ in sizes]
    bc.plan_digit_bias(6)
    assert bc.plan(40000, 40000)["bits"] == 20
    bc.stats_reset()
    out, worst = bc.mul_batch(pairs, adds=adds, cls=[len(x) + len(y) + 1 for x, y in pairs])
    st = bc.stats()
    assert st["retries"] >= 2, st
    assert worst < 0.25
    for (x, y), d, c in zip(pairs, adds, out):
        assert words_to_int(c) == words_to_int(x) * words_to_int(y) + words_to_int(d)
    x, y = pairs[3]
    c, info = bc.muladd(x, y, adds[3])
    assert info["retries"] >= 1 and words_to_int(c) == words_to_int(x) * words_to_int(y) + words_to_int(adds[3])


def test_margin_retry_e_pipeline(bc, golden, no_bias):
    bc.plan_digit_bias(6)
    bc.stats_reset()
    txt, _ = bc.e_digits(10 ** 5)
    st = bc.stats()
    assert st["retries"] > 0, st
    assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"]["100000"]


def test_digit_bias_argument_checked(bc):
    with pytest.raises(bc.GpuError):
        bc.plan_digit_bias(-1)
    with pytest.raises(bc.GpuError):
        bc.plan_digit_bias(9)
