"""GPU parity of the multiply path (_fft_mul, bigfloat_mul, the device C ABI)
against exact integer products, the golden fixtures and the oracle. Integer
results must be bit-exact; every product's rounding margin must be < 0.25."""
import numpy as np
import pytest

from helpers import exact_product, sha, words_to_int, int_to_words
from bluecrunch_amd.workload import operands, product_fingerprint_ok

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (2, 3), (49, 50), (50, 50), (1, 150), (99, 1), (100, 100), (127, 129), (300, 1000), (1000, 7),
          (2000, 2000), (4095, 4097), (3000, 9000), (12345, 777), (16384, 16384), (40000, 40000), (100000, 70000),
          (131072, 131072), (1, 100000), (300000, 299999)]


@pytest.mark.parametrize("al,bl", SHAPES)
def test_mul_exact(bc, al, bl):
    a, b = operands(1000 + al + 7 * bl, al, bl)
    c, info = bc.mul(a, b)
    assert np.array_equal(c, exact_product(a, b)), info
    assert info["margin"] < 0.25


def test_mul_random_shapes(bc):
    # 80 seeded operand shapes, lengths log-uniform in [1, 3e5] words (every
    # planner digit size, 2^k and 3 2^k lengths, one- and two-pass plans,
    # balanced and lopsided operands, either side of the lower-half bound),
    # each product exact with its margin below 0.25
    rng = np.random.default_rng(2024)
    plans = set()
    for i in range(80):
        al, bl = (int(np.exp(rng.uniform(0, np.log(3e5)))) for _ in range(2))
        a, b = operands(50000 + i, al, bl)
        c, info = bc.mul(a, b)
        assert np.array_equal(c, exact_product(a, b)), (al, bl, info)
        assert info["margin"] < 0.25, (al, bl, info)
        plans.add((info["points"], info["bits"], info["passes"]))
    assert len(plans) > 30


def test_random_shapes_muladd_square_alias(bc):
    # 40 seeded shapes up to 1e5 words: a*b + d truncated to CL words (CL
    # below, at and above AL + BL; addends shorter and longer than the
    # product), the square of a, and the product written over A in place
    import ctypes
    rng = np.random.default_rng(77)
    for i in range(40):
        al, bl = (int(np.exp(rng.uniform(0, np.log(1e5)))) for _ in range(2))
        a, b = operands(60000 + i, al, bl)
        dl = int(rng.integers(0, al + bl + 4))
        d = operands(70000 + i, max(dl, 1), 1)[0][:dl]
        cl = max(1, al + bl + int(rng.integers(-3, 3)))
        c, info = bc.muladd(a, b, d, cl)
        want = (words_to_int(a) * words_to_int(b) + words_to_int(d)) % (1 << (32 * cl))
        assert words_to_int(c) == want, ("muladd", al, bl, dl, cl, info)
        s, info = bc.mul(a, a)
        assert np.array_equal(s, exact_product(a, a)), ("square", al, info)
        dA = bc.DeviceWords(al + bl)
        bc.check(bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, 4 * al))
        dB = bc.DeviceWords.from_host(b)
        info = bc.mul_dev(dA.ptr, al + bl, dA.ptr, al, dB.ptr, bl)
        assert np.array_equal(dA.to_host(), exact_product(a, b)), ("alias", al, bl, info)


@pytest.mark.parametrize("al", [1, 60, 100, 1024, 5000, 65536, 262144])
def test_square_exact(bc, al):
    a, _ = operands(al, al, 1)
    c, info = bc.mul(a, a)
    assert np.array_equal(c, exact_product(a, a))
    if al + al >= 100:
        assert info["square"] == 1


@pytest.mark.parametrize("al", [100, 4096, 65536, 1 << 20])
def test_all_ones_adversarial(bc, al):
    # maximal digits: the worst magnitude for the unsigned digit FFT
    a = np.full(al, 0xffffffff, dtype=np.uint32)
    c, info = bc.mul(a, a)
    x = (1 << (32 * al)) - 1
    assert words_to_int(c) == x * x
    assert info["margin"] < 0.25


def test_golden_product_hashes(bc, golden):
    for p in golden["products"]:
        a, b = operands(p["seed"], p["al"], p["bl"])
        c, _ = bc.mul(a, b)
        assert sha(c) == p["sha256"], p
        s, _ = bc.mul(a, a)
        assert sha(s) == p["square_sha256"], p


@pytest.mark.parametrize("cl_delta", [-50, -1, 1, 5])
def test_output_length_other_than_al_plus_bl(bc, cl_delta):
    # _fft_mul writes exactly CL words: the low CL words of the product
    a, b = operands(77, 600, 500)
    cl = 1100 + cl_delta
    c, _ = bc.mul(a, b, cl=cl)
    assert np.array_equal(c, exact_product(a, b, cl))


def test_fft_mul_entry_point(bc):
    a, b = operands(5, 5000, 4000)
    c, bits = bc.fft_mul(a, b)
    assert np.array_equal(c, exact_product(a, b)) and 4 <= bits <= 20


@pytest.mark.parametrize("al,bl", [(3000001, 2500003), (2097152, 2097153), (4194303, 1)])
def test_fft_mul_host_in_place_across_staging_chunks(bc, oracle, al, bl):
    # operands and product larger than one 8 MiB staging chunk, lengths not a
    # multiple of it; CT is AT's buffer (in-place, as bigfloat_mul calls it)
    a, b = operands(11, al, bl)
    want, _, _ = oracle.fft_mul(a, b)
    buf = np.zeros(al + bl, dtype=np.uint32)
    buf[:al] = a
    bits = bc.lib()._fft_mul(bc._u32(buf), al + bl, bc._u32(buf), al, bc._u32(b), bl)
    assert 4 <= bits <= 20
    assert np.array_equal(buf, want)


def test_basecase_and_karatzuba_entry_points(bc):
    import ctypes
    for al, bl in [(1, 1), (3, 17), (49, 50), (200, 300)]:
        a, b = operands(al + bl, al, bl)
        c = bc.basecase(a, b)
        assert np.array_equal(c, exact_product(a, b))
        out = np.zeros(al + bl, dtype=np.uint32)
        bc.lib()._karatzuba_mul(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), al + bl,
                                a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), al,
                                b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), bl)
        assert np.array_equal(out, exact_product(a, b))


def test_bigfloat_mul_regression_square(bc, golden):
    # test/test_regression.c:40-67 -- in-place a = a*a
    g = golden["regression_square"]
    a = (np.array(g["a"], np.uint32), 0, 1)
    w, e, s = bc.bigfloat_mul(a, a, 0, inplace="a")
    assert list(map(int, w)) == g["expected"] and e == 0 and s == 1


def test_bigfloat_exp_vector(bc, golden):
    # test/test_bigfloat_exp.c -- 0x1234^0xd7 by square-and-multiply
    g = golden["bigfloat_exp"]
    target = (np.array([1], np.uint32), 0, 1)
    cp = (np.array([g["base"]], np.uint32), 0, 1)
    e = g["exp"]
    while e:
        if e & 1:
            target = bc.bigfloat_mul(target, cp, 0)
        cp = bc.bigfloat_mul(cp, cp, 0)
        e >>= 1
    w, ex, s = target
    assert len(w) == g["len"] and ex == 0 and s == 1 and words_to_int(w) == int(g["hex"], 16)


def test_bigfloat_exp_through_fft_sizes(bc):
    # square-and-multiply far into the FFT range, checked against Python ints
    base, e = 0x9e3779b9, 5000
    target = (np.array([1], np.uint32), 0, 1)
    cp = (np.array([base], np.uint32), 0, 1)
    k = e
    while k:
        if k & 1:
            target = bc.bigfloat_mul(target, cp, 0)
        cp = bc.bigfloat_mul(cp, cp, 0)
        k >>= 1
    assert words_to_int(target[0]) == base ** e


@pytest.mark.parametrize("p", [0, 1, 50, 120, 2000])
def test_bigfloat_mul_matches_oracle(bc, oracle, p):
    a, b = operands(31 + p, 700, 1300)
    for sa, sb in [(1, 1), (0, 1), (0, 0)]:
        A, B = (a, 4, sa), (b, -9, sb)
        got = bc.bigfloat_mul(A, B, p)
        want = oracle.bigfloat_mul(A, B, p)
        assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]


def test_bigfloat_mul_zero(bc):
    a, _ = operands(3, 10, 1)
    w, e, s = bc.bigfloat_mul((np.zeros(0, np.uint32), 3, 0), (a, 0, 1), 0)
    assert len(w) == 0 and e == 0 and s == 1


def test_device_abi_alias_and_repeat(bc):
    # C may alias A (in-place), repeated calls reuse buffers
    a, b = operands(8, 50000, 50000)
    want = exact_product(a, b)
    dA = bc.DeviceWords(100000)
    import ctypes
    bc.check(bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, 4 * a.size))
    dB = bc.DeviceWords.from_host(b)
    for _ in range(2):
        info = bc.mul_dev(dA.ptr, 100000, dA.ptr, 50000, dB.ptr, 50000)
        assert np.array_equal(dA.to_host(), want)
        bc.check(bc.lib().bcg_memcpy_h2d(dA.ptr, a.ctypes.data, 4 * a.size))


@pytest.mark.parametrize("words", [1 << 20, 1 << 22])
def test_large_against_oracle(bc, oracle, words):
    a, b = operands(1, words, words)
    c, info = bc.mul(a, b)
    o, _, _ = oracle.fft_mul(a, b)
    assert np.array_equal(c, o)
    assert info["margin"] < 0.25


def test_benchmark_size_fingerprint(bc):
    # 2^24 x 2^24 words (the benchmark workload): size-independent modular check
    w = 1 << 24
    a, b = operands(1, w, w)
    c, info = bc.mul(a, b)
    assert product_fingerprint_ok(a, b, c)
    assert info["bits"] == 8 and info["points"] == 1 << 26 and info["margin"] < 0.25
    s, sinfo = bc.mul(a, a)
    assert product_fingerprint_ok(a, a, s) and sinfo["margin"] < 0.25


@pytest.mark.parametrize("bits", [8, 10, 12, 14, 16])
def test_forced_digit_sizes(bc, bits):
    a, b = operands(bits, 20000, 20000)
    dA, dB, dC = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b), bc.DeviceWords(40000)
    info = bc.mul_dev(dC.ptr, 40000, dA.ptr, 20000, dB.ptr, 20000, bits=bits)
    assert info["bits"] <= bits
    assert np.array_equal(dC.to_host(), exact_product(a, b))


@pytest.mark.parametrize("bits", [5, 6, 7, 9, 11, 13])
def test_forced_digit_sizes_ragged(bc, bits):
    # odd lengths leave the upper half of the packing pass partly filled and,
    # with odd digit sizes, digits straddle words (two-word loads) up to the
    # operands' last word, and group starts are not word-aligned in the carry
    a, b = operands(100 + bits, 20011, 19993)
    dA, dB, dC = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b), bc.DeviceWords(40004)
    info = bc.mul_dev(dC.ptr, 40004, dA.ptr, 20011, dB.ptr, 19993, bits=bits)
    assert info["bits"] <= bits and info["passes"] > 1
    assert np.array_equal(dC.to_host(), exact_product(a, b, 40004))


@pytest.mark.parametrize("words", [20000, 100000, 300000, 1 << 22])
@pytest.mark.parametrize("side", ["lower", "upper", "square"])
def test_packing_lower_half_boundary(bc, words, side):
    # The packing pass skips the upper digit half (digits N .. 2N - 1) when
    # both operands fit in h = N b / 32 words. Products on each side of h for
    # the plan of (words, words): (h, h - 9) and the square of h words skip
    # it, (h + 1, h - 9) does not. The plan must stay the same so the case
    # lands on the intended side. The three-pass size is checked by fingerprint.
    p = bc.plan(words, words)
    h = p["points"] * p["bits"] // 32
    al, bl = {"lower": (h, h - 9), "upper": (h + 1, h - 9), "square": (h, h)}[side]
    assert bc.plan(al, bl) == p
    a, b = operands(7 * words + len(side), al, bl)
    if side == "square":
        b = a
    cl = al + bl
    dA, dC = bc.DeviceWords.from_host(a), bc.DeviceWords(cl)
    dB = dA if side == "square" else bc.DeviceWords.from_host(b)
    info = bc.mul_dev(dC.ptr, cl, dA.ptr, al, dB.ptr, bl)
    assert (info["points"], info["bits"], info["passes"]) == (p["points"], p["bits"], p["passes"])
    assert info["margin"] < 0.25
    c = dC.to_host()
    if p["passes"] == 3:
        assert product_fingerprint_ok(a, b, c)
    else:
        assert np.array_equal(c, exact_product(a, b))


@pytest.mark.parametrize("words", [20011, 300007])
def test_muladd_group_carry(bc, words):
    # a*b + d through the group-sum carry (multi-pass sizes): the addend's
    # words enter every carry block they reach, a carry runs out of its top
    a, b = operands(words, words, words - 17)
    d = np.full(words + 5, 0xffffffff, dtype=np.uint32)
    c, info = bc.muladd(a, b, d)
    want = words_to_int(a) * words_to_int(b) + words_to_int(d)
    assert words_to_int(c) == want % (1 << (32 * c.size))
    assert info["margin"] < 0.25


@pytest.mark.parametrize("words,stride", [(9000, 1), (100000, 1), (700000, 5)])
def test_every_pass_split_is_exact(bc, words, stride):
    # every factorisation the autotuner may pick must give the same exact
    # product (at the largest size every fifth split plus the default one)
    a, b = operands(words, words, words)
    want = exact_product(a, b)
    want_sq = exact_product(a, a)
    n = bc.plan(words, words)["points"]
    cands = bc.plan_candidates(n)
    assert cands
    picks = sorted(set(range(0, len(cands), stride)) | {bc.plan_current(n)})
    try:
        for i in picks:
            bc.plan_force(n, i)
            c, info = bc.mul(a, b)
            assert np.array_equal(c, want), (i, cands[i])
            s, _ = bc.mul(a, a)
            assert np.array_equal(s, want_sq), (i, cands[i])
    finally:
        bc.plan_force(n, -1)


@pytest.mark.parametrize("words,group_bytes,exact", [(500000, [1, 5 << 20], True), (1 << 21, [24 << 20], False)])
def test_chunked_middle_passes(bc, words, group_bytes, exact):
    # three-pass splits run their middle passes group by group of blocks (the
    # points of a group stay in the memory-side cache between launches). With
    # small groups (one block each, or several with a short last group) every
    # three-pass split must give the words of whole-transform passes, for
    # products and squares, and those words the exact product (checked by
    # Python integers at 500000 words, by fingerprint at 2^21 words)
    a, b = operands(77 + words, words, words)
    n = bc.plan(words, words)["points"]
    cands = bc.plan_candidates(n)
    picks = [i for i, (P, L, C) in enumerate(cands) if P == 3]
    assert picks
    want = (exact_product(a, b), exact_product(a, a)) if exact else None
    chunked = 0
    try:
        for i in picks:
            bc.plan_force(n, i)
            bc.plan_chunk(0)
            c0, info0 = bc.mul(a, b)
            s0, _ = bc.mul(a, a)
            assert info0["chunks"] == 0
            if exact:
                assert np.array_equal(c0, want[0]), (i, cands[i])
                assert np.array_equal(s0, want[1]), (i, cands[i])
            else:
                assert product_fingerprint_ok(a, b, c0) and product_fingerprint_ok(a, a, s0), (i, cands[i])
            for gb, streams in [(g, k) for g in group_bytes for k in (1, 2, 3)]:
                bc.plan_chunk(gb, streams)
                c, info = bc.mul(a, b)
                s, sinfo = bc.mul(a, a)
                assert np.array_equal(c, c0), (i, cands[i], gb, streams, info)
                assert np.array_equal(s, s0), (i, cands[i], gb, streams, sinfo)
                assert info["margin"] < 0.25 and sinfo["margin"] < 0.25
                chunked += info["chunks"] > 1 and sinfo["chunks"] > 1
    finally:
        bc.plan_chunk(128 << 20)
        bc.plan_force(n, -1)
    assert chunked


def test_chunked_groups_at_2p24_words(bc):
    # the 2^24-word product with the 256 x 256 x 1024 split (2^26 points)
    # runs its middle passes in groups at the default group size; its words
    # equal those of whole-transform passes and pass the fingerprint check
    a, b = operands(1, 1 << 24, 1 << 24)
    n = bc.plan(1 << 24, 1 << 24)["points"]
    cands = bc.plan_candidates(n)
    idx = [i for i, (P, L, C) in enumerate(cands) if L == [256, 256, 1024] and C == [16, 16, 1]]
    assert idx, cands
    try:
        bc.plan_force(n, idx[0])
        c, info = bc.mul(a, b)
        assert info["passes"] == 3 and info["chunks"] > 1, info
        assert product_fingerprint_ok(a, b, c)
        bc.plan_chunk(0)
        c0, info0 = bc.mul(a, b)
    finally:
        bc.plan_chunk(128 << 20)
        bc.plan_force(n, -1)
    assert info0["chunks"] == 0
    assert np.array_equal(c, c0)


@pytest.mark.parametrize("extra", [3, (1 << 19) + 5])
def test_one_pass_long_outputs(bc, extra):
    # one-pass products whose output runs past the product: zero words, or a
    # long addend; from 2^19 output words the workgroup carry takes its
    # 64-bit form
    a, b = operands(11, 3000, 2000)
    cl = a.size + b.size + extra
    c, info = bc.mul(a, b, cl=cl)
    assert info["passes"] == 1, info
    want = exact_product(a, b)
    assert np.array_equal(c[:want.size], want) and not c[want.size:].any()
    d = np.random.default_rng(5).integers(0, 1 << 32, cl, dtype=np.uint64).astype(np.uint32)
    d[-9:] = 0xffffffff  # a carry through the addend's top words
    c2, info2 = bc.muladd(a, b, d, cl=cl)
    assert info2["passes"] == 1, info2
    x = (words_to_int(want) + words_to_int(d)) % (1 << (32 * cl))
    assert words_to_int(c2) == x


def test_autotune_keeps_exactness(bc):
    chosen, times = bc.autotune(300000, 300000, reps=1)
    assert 0 <= chosen < len(times) and all(t > 0 for t in times)
    a, b = operands(4, 300000, 300000)
    c, _ = bc.mul(a, b)
    assert product_fingerprint_ok(a, b, c)


@pytest.mark.parametrize("k", [1000, 5000, 70000, 1 << 18])
def test_carry_through_all_ones_blocks(bc, k):
    # (2^(32k) - 1)(2^(32k) + 1) + 1 = 2^(64k): before the final +1 every word
    # is 0xffffffff, so the carry enters whole carry blocks that propagate
    # (they become zero) and stops in the top block after a run of ones
    a = np.full(k, 0xffffffff, dtype=np.uint32)
    b = np.zeros(k + 1, dtype=np.uint32)
    b[0] = b[k] = 1
    c, info = bc.muladd(a, b, np.array([1], dtype=np.uint32), 2 * k + 2)
    want = np.zeros(2 * k + 2, dtype=np.uint32)
    want[2 * k] = 1
    assert np.array_equal(c, want), int(np.flatnonzero(c != want)[0])
    assert info["margin"] < 0.25
    # (2^(32k) - 2)(2^(32k) + 1) + 2 = 2^(64k) - 2^(32k): the carry out of the
    # low word runs through the k - 1 words of ones above it, stops at word k
    a[0] = 0xfffffffe
    c, _ = bc.muladd(a, b, np.array([2], dtype=np.uint32), 2 * k + 2)
    x = ((1 << (32 * k)) - 2) * ((1 << (32 * k)) + 1) + 2
    assert words_to_int(c) == x


def test_batched_carry_through_all_ones_blocks(bc):
    pairs, adds, want = [], [], []
    for k in (300, 2000, 9000):
        a = np.full(k, 0xffffffff, dtype=np.uint32)
        b = np.zeros(k + 1, dtype=np.uint32)
        b[0] = b[k] = 1
        pairs.append((a, b))
        adds.append(np.array([1], dtype=np.uint32))
        w = np.zeros(2 * k + 1, dtype=np.uint32)
        w[2 * k] = 1
        want.append(w)
    out, margin = bc.mul_batch(pairs, adds=adds)
    for c, w in zip(out, want):
        assert np.array_equal(c, w)
    assert margin < 0.25


def test_running_stats(bc):
    bc.stats_reset()
    assert bc.stats()["products"] == 0
    a, b = operands(5, 5000, 7000)
    _, info = bc.mul(a, b)
    st = bc.stats()
    assert st["products"] == 1
    assert st["max_points"] == info["points"] and st["bits_at_max"] == info["bits"]
    assert st["margin_at_max"] == info["margin"] and st["worst_margin"] == info["margin"]
    # a batch counts each FFT product; the tiny ones (schoolbook) are not FFT products
    pairs = [operands(10 + k, 300 + 50 * k, 200) for k in range(5)] + [operands(99, 10, 20)]
    _, worst = bc.mul_batch(pairs)
    st = bc.stats()
    assert st["products"] == 6
    assert st["worst_margin"] >= worst
    bc.stats_reset()
    assert bc.stats()["max_points"] == 0
