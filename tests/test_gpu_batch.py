"""Batched uneven products (persistent one-workgroup class kernels) against
exact integer products."""
import numpy as np
import pytest

from helpers import exact_product
from bluecrunch_amd.workload import SplitMix64, batch_lengths, operand

pytestmark = pytest.mark.gpu


def make_batch(seed, count, lo=16, hi=16384):
    lens = batch_lengths(seed, count, lo, hi)
    rng = SplitMix64(seed + 1)
    return [(operand(rng, int(x)), operand(rng, int(y))) for x, y in lens]


def test_batch_small_uneven(bc):
    pairs = make_batch(7, 200, 16, 4096)
    out, worst = bc.mul_batch(pairs)
    for (a, b), c in zip(pairs, out):
        assert np.array_equal(c, exact_product(a, b))
    assert worst < 0.25


def test_batch_mixed_with_squares_and_tiny(bc):
    rng = SplitMix64(99)
    pairs = []
    for n in [1, 2, 10, 49, 50, 60, 200, 1000, 3000, 8000, 20000]:
        a = operand(rng, n)
        pairs.append((a, a))                  # square (same array object)
        pairs.append((a, operand(rng, max(1, n // 3))))
    out, worst = bc.mul_batch(pairs)
    for (a, b), c in zip(pairs, out):
        assert np.array_equal(c, exact_product(a, b))


def test_batch_benchmark_shape(bc):
    # the batched workload: log-uniform operand lengths in [16, 16384], seed 7
    pairs = make_batch(7, 512)
    out, worst = bc.mul_batch(pairs)
    bad = [i for i, ((a, b), c) in enumerate(zip(pairs, out)) if not np.array_equal(c, exact_product(a, b))]
    assert not bad
    assert worst < 0.25


def test_batch_multipass_classes(bc):
    # many two-pass products sharing transform lengths, plus a few odd sizes
    rng = SplitMix64(5)
    pairs = [(operand(rng, 9000 + 13 * i), operand(rng, 7000 + 29 * i)) for i in range(40)]
    pairs += [(operand(rng, 30000), operand(rng, 100)), (operand(rng, 16384), operand(rng, 16384))]
    out, worst = bc.mul_batch(pairs)
    for (a, b), c in zip(pairs, out):
        assert np.array_equal(c, exact_product(a, b))
    assert worst < 0.25


def test_muladd_single_all_paths(bc):
    from helpers import words_to_int, int_to_words
    rng = SplitMix64(11)
    for al, bl, dl in [(3, 4, 9), (40, 50, 100), (300, 200, 600), (5000, 3000, 8000), (40000, 30000, 70000)]:
        a, b, d = operand(rng, al), operand(rng, bl), operand(rng, dl)
        cl = max(al + bl, dl) + 1
        c, _ = bc.muladd(a, b, d, cl)
        want = (words_to_int(a) * words_to_int(b) + words_to_int(d)) % (1 << (32 * cl))
        assert words_to_int(c) == want, (al, bl, dl)


def test_muladd_batch_binary_splitting_shape(bc):
    # P = P0*Q1 + P1 for many nodes of mixed sizes, tiny to two-pass
    from helpers import words_to_int, int_to_words
    rng = SplitMix64(12)
    pairs, adds, cls = [], [], []
    for n in [1, 1, 2, 3, 7, 20, 49, 60, 128, 700, 2500, 9000, 20000] * 3:
        a, b, d = operand(rng, n), operand(rng, n + 1), operand(rng, n + 1)
        pairs.append((a, b)); adds.append(d); cls.append(2 * n + 2)
    out, worst = bc.mul_batch(pairs, cls=cls, adds=adds)
    for (a, b), d, c, cl in zip(pairs, adds, out, cls):
        want = (words_to_int(a) * words_to_int(b) + words_to_int(d)) % (1 << (32 * cl))
        assert words_to_int(c) == want
