"""Shared test helpers: exact integer arithmetic on word arrays, glibc rand()."""
import ctypes
import hashlib

import numpy as np

from bluecrunch_amd.workload import int_to_words, words_to_int  # noqa: F401


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()


def exact_product(a, b, cl=None):
    n = len(a) + len(b) if cl is None else cl
    p = words_to_int(a) * words_to_int(b)
    return int_to_words(p % (1 << (32 * n)), n)


class GlibcRand:
    """glibc srand/rand, the generator test/test_conversion.c uses."""

    def __init__(self, seed):
        self.libc = ctypes.CDLL("libc.so.6")
        self.libc.srand.argtypes = [ctypes.c_uint]
        self.libc.rand.restype = ctypes.c_int
        self.libc.srand(seed)

    def __call__(self):
        return self.libc.rand()


def bitrev(k, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (k & 1)
        k >>= 1
    return r


def reference_order(n):
    """Position -> frequency map of the reference fft_forward output
    (bit-reversed, radix-3 digit innermost for 3*2^k)."""
    three = n % 3 == 0
    p2 = n // 3 if three else n
    lg = p2.bit_length() - 1
    freq = np.zeros(n, dtype=np.int64)
    for pos in range(n):
        if three:
            freq[pos] = bitrev(pos // 3, lg) + (pos % 3) * p2
        else:
            freq[pos] = bitrev(pos, lg)
    return freq


def brute_dft(x, sign=+1):
    """O(N^2) DFT in long double: X_k = sum_j x_j e^{sign 2 pi i jk/N}."""
    n = len(x)
    j = np.arange(n, dtype=np.longdouble)
    xr = np.asarray(x.real, dtype=np.longdouble)
    xi = np.asarray(x.imag, dtype=np.longdouble)
    out = np.zeros(n, dtype=np.complex128)
    twopi = np.longdouble("6.283185307179586476925286766559005768")
    for k in range(n):
        ang = sign * twopi * ((j * k) % n) / n
        c, s = np.cos(ang), np.sin(ang)
        re = (xr * c - xi * s).sum()
        im = (xr * s + xi * c).sum()
        out[k] = complex(float(re), float(im))
    return out
