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


def test_muladd_batch_random(bc):
    # 300 seeded products, lengths log-uniform in [1, 40000] words, addends of
    # 0 to CL + 2 words, outputs from CL = AL + BL - 2 to AL + BL + 2 words: tiny,
    # one-pass and multi-pass classes in one call, each product exact
    from helpers import words_to_int
    rng = np.random.default_rng(31)
    sm = SplitMix64(31)
    pairs, adds, cls = [], [], []
    for _ in range(300):
        al, bl = (int(np.exp(rng.uniform(0, np.log(40000)))) for _ in range(2))
        cl = max(1, al + bl + int(rng.integers(-2, 3)))
        pairs.append((operand(sm, al), operand(sm, bl)))
        adds.append(operand(sm, int(rng.integers(0, cl + 3))))
        cls.append(cl)
    out, worst = bc.mul_batch(pairs, cls=cls, adds=adds)
    for i, ((a, b), d, c, cl) in enumerate(zip(pairs, adds, out, cls)):
        want = (words_to_int(a) * words_to_int(b) + words_to_int(d)) % (1 << (32 * cl))
        assert words_to_int(c) == want, (i, len(a), len(b), len(d), cl)
    assert worst < 0.25


def _device_batch(bc, pairs, adds=None):
    al = np.array([len(a) for a, _ in pairs], dtype=np.uint64)
    bl = np.array([len(b) for _, b in pairs], dtype=np.uint64)
    cl = al + bl
    offs = lambda n: np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)  # noqa: E731
    dA = bc.DeviceWords.from_host(np.concatenate([a for a, _ in pairs]))
    dB = bc.DeviceWords.from_host(np.concatenate([b for _, b in pairs]))
    dC = bc.DeviceWords(int(cl.sum()))
    extra = {}
    if adds is not None:
        dl = np.array([len(d) for d in adds], dtype=np.uint64)
        extra = dict(dD=bc.DeviceWords.from_host(np.concatenate(adds)), d_off=offs(dl), dl=dl)
    return dA, dB, dC, (offs(cl), cl, offs(al), al, offs(bl), bl), extra


def test_batch_graph_replays(bc):
    # one hipGraph for tiny, one-pass and multi-pass products (and an addend),
    # replayed three times with eager library work and the output cleared in
    # between; every replay exact
    rng = SplitMix64(123)
    pairs = [(operand(rng, n), operand(rng, m)) for n, m in
             [(3, 5), (20, 30), (200, 150), (1000, 3000), (4000, 4000), (16384, 9000), (30000, 20000), (7, 90)]]
    adds = [operand(rng, 5) for _ in pairs]
    dA, dB, dC, (c_off, cl, a_off, al, b_off, bl), extra = _device_batch(bc, pairs, adds)
    want = [exact_product(a, b) for a, b in pairs]
    for (d, w, n) in zip(adds, want, cl):
        x = int.from_bytes(w.astype("<u4").tobytes(), "little") + int.from_bytes(d.astype("<u4").tobytes(), "little")
        w[:] = np.frombuffer((x % (1 << (32 * int(n)))).to_bytes(4 * int(n), "little"), dtype="<u4")
    batch = bc.Batch(dC, dA, dB, c_off, cl, a_off, al, b_off, bl, **extra)
    try:
        for rep in range(3):
            dC.zero()
            worst = batch.run()
            C = dC.to_host()
            for i, w in enumerate(want):
                assert np.array_equal(C[int(c_off[i]):int(c_off[i] + cl[i])], w), (rep, i)
            assert worst < 0.25
            # eager work between replays (other products through the same library)
            x, y = operand(rng, 5000), operand(rng, 7000)
            assert np.array_equal(bc.mul(x, y)[0], exact_product(x, y))
    finally:
        batch.close()
