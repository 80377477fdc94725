"""GPU parity of the fft.h entry points against the oracle and the
reference's known-answer vectors (tolerance 1e-10 absolute, as
test/test_fft.c:10)."""
import numpy as np
import pytest

from helpers import GlibcRand, brute_dft, reference_order, words_to_int

pytestmark = pytest.mark.gpu

LENGTHS = [2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 512, 768, 1024, 2048, 3072, 4096, 6144, 8192,
           12288, 16384, 1 << 15, 3 << 14, 1 << 17, 3 << 16, 1 << 18, 3 << 17, 1 << 19, 3 << 18, 1 << 20, 3 << 19,
           1 << 21, 3 << 20, 1 << 22]


def cplx(pairs):
    return np.array([complex(a, b) for a, b in pairs], dtype=np.complex128)


def test_fft4_reference_vector(bc, golden):
    g = golden["fft4"]
    z = cplx(g["input"])
    bc.fft_forward(z)
    assert np.abs(z - cplx(g["forward"])).max() < 1e-10
    bc.fft_inverse(z)
    assert np.abs(z - cplx(g["inverse"])).max() < 1e-10


@pytest.mark.parametrize("n", LENGTHS)
def test_forward_inverse_match_oracle(bc, oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = x.copy(); bc.fft_forward(z)
    o = x.copy(); oracle.fft_forward(o)
    scale = max(1.0, np.abs(o).max())
    assert np.abs(z - o).max() <= 1e-13 * scale * np.log2(n + 1)
    back = z.copy(); bc.fft_inverse(back)
    ob = o.copy(); oracle.fft_inverse(ob)
    assert np.abs(back - ob).max() <= 1e-12 * n
    assert np.abs(back / n - x).max() < 1e-12


@pytest.mark.parametrize("n", [64, 96, 192, 1024, 1536])
def test_forward_brute_force(bc, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = x.copy(); bc.fft_forward(z)
    want = brute_dft(x, +1)[reference_order(n)]
    assert np.abs(z - want).max() < 1e-10


def test_baileys_forward_is_the_forward_transform(bc):
    n = 1 << 16
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = x.copy(); bc.baileys_forward(a)
    b = x.copy(); bc.fft_forward(b)
    assert np.array_equal(a, b)


def test_pointwise(bc):
    rng = np.random.default_rng(2)
    a = rng.standard_normal(10000) + 1j * rng.standard_normal(10000)
    b = rng.standard_normal(10000) + 1j * rng.standard_normal(10000)
    want = a * b
    bc.fft_pointwise(a, b)
    assert np.abs(a - want).max() < 1e-13


def test_int_to_fft_vectors(bc, golden):
    for case in golden["int_to_fft"]:
        z, n = bc.int_to_fft(4, case["words"], case["bits"])
        pts = case["points"]
        assert n == len(pts)
        assert np.array_equal(z[:n].real, np.array(pts, float)) and not z[n:].any() and not z.imag.any()


def test_fft_to_int_vectors(bc, golden):
    for case in golden["fft_to_int"]:
        z = np.zeros(16, complex)
        z[:len(case["points"])] = np.array(case["points"], float) * 16
        w = bc.fft_to_int(z, case["length"], len(case["words"]), case["bits"])
        assert list(map(int, w)) == case["words"]


def test_conversion_round_trip_glibc(bc, golden):
    # test/test_conversion.c:18-45 (first 300 runs of the same glibc stream)
    g = golden["conversion"]
    rnd = GlibcRand(g["seed"])
    for _ in range(300):
        values = np.array([rnd() for _ in range(g["words"])], dtype=np.uint32)
        ln = rnd() % 64
        bpp = 7 + rnd() % (22 - 7)
        z, n = bc.int_to_fft(g["buffer"], values[:ln], bpp)
        out = bc.fft_to_int(z, 1, ln, bpp)
        assert np.array_equal(out, values[:ln]), (ln, bpp)


@pytest.mark.parametrize("bits", [7, 8, 11, 16, 22])
def test_conversions_match_oracle(bc, oracle, bits):
    rng = np.random.default_rng(bits)
    w = rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)
    z, n = bc.int_to_fft(32768, w, bits)
    zo, no = oracle.int_to_fft(32768, w, bits)
    assert n == no and np.array_equal(z[:32768], zo[:32768])
    # fft_to_int with carries: points are coefficients times length
    r = rng.integers(0, 1 << 40, n).astype(float)
    zz = (r + rng.uniform(-0.2, 0.2, n)) * 64 + 0j
    got, m = bc.fft_to_int_margin(zz, 64, 5000, bits)
    want, mo = oracle.fft_to_int(zz, 64, 5000, bits)
    assert np.array_equal(got, want)
    assert abs(m - mo) < 1e-6


def test_fft_length(bc, oracle):
    for n in range(0, 3000, 7):
        assert bc.fft_length(n) == oracle.fft_length(n)


def test_tft_reference_vectors(bc, golden):
    # test/test_tft.c -- forward and inverse known answers (tolerance 1e-6 as there)
    for case in golden["tft_forward"]:
        z = cplx(case["input"])
        bc.tft_forward(z, case["len"], case["k"])
        assert np.abs(z - cplx(case["expected"])).max() < 1e-6
    for case in golden["tft_inverse"]:
        z = cplx(case["input"])
        bc.tft_inverse(z, case["len"], case["k"])
        assert np.abs(z - np.array(case["expected"], complex)).max() < 1e-6, case["len"]


@pytest.mark.parametrize("k,ln", [(4, 9), (5, 17), (6, 33), (8, 200), (10, 513), (10, 1000), (12, 2049),
                                  (12, 3000), (16, 40000), (16, 65535)])
def test_tft_round_trip_and_definition(bc, k, ln):
    n = 1 << k
    rng = np.random.default_rng(k * 1000 + ln)
    x = np.zeros(n, complex)
    x[:ln] = rng.standard_normal(ln) + 1j * rng.standard_normal(ln)
    z = x.copy()
    bc.tft_forward(z, ln, k)
    full = x.copy(); bc.fft_forward(full)
    assert np.abs(z[:ln] - full[:ln]).max() < 1e-9 * n and not z[ln:].any()
    bc.tft_inverse(z, ln, k)
    assert np.abs(z[:ln] - n * x[:ln]).max() < 1e-7 * n and not z[ln:].any()


def test_twiddle_tables_through_the_library(bc):
    # fft/table.c:211-234 through the library's exported globals: the slot
    # count of the latest fft_ensure_table call, table_select's -1 past it, and
    # table values e^{+2 pi i c/L} (c < L/2) in slot _table_select(L/2)
    import ctypes
    import math
    L = bc.lib()
    size = ctypes.c_int.in_dll(L, "twiddle_table_size")
    tabs = (ctypes.c_void_p * 125).in_dll(L, "twiddle_table")
    L.fft_free_table.restype = None
    L.fft_ensure_table(3 << 12)
    # k stops at 8192 (2k = 16384 > 12288): slots 0 .. _table_select(8192) = 24
    assert size.value == 25
    assert L.table_select(4096) == 22 and L.table_select(6144) == 23 and L.table_select(8192) == 24
    for Lc, slot in [(4, 0), (6, 1), (64, 8), (96, 9), (8192, 22), (12288, 23)]:
        assert L.table_select(Lc // 2) == slot
        t = np.ctypeslib.as_array(ctypes.cast(tabs[slot], ctypes.POINTER(ctypes.c_double)), shape=(Lc,))
        z = t[0::2] + 1j * t[1::2]
        c = np.arange(Lc // 2)
        want = np.array([complex(math.cos(2 * math.pi * k / Lc), math.sin(2 * math.pi * k / Lc)) for k in c])
        assert np.max(np.abs(z - want)) < 4e-16 * max(1.0, math.log2(Lc)), Lc
        assert z[0] == 1
    # a smaller call resets the size (fft/table.c:227), so bigger slots read as absent
    L.fft_ensure_table(100)
    assert size.value == 11
    assert L.table_select(64) == 10 and L.table_select(4096) == ctypes.c_size_t(-1).value
    L.fft_free_table()
    assert size.value == 0 and all(tabs[i] is None for i in range(125))
