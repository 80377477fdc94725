"""Seeded synthetic operands for the benchmark and the tests.

Words are drawn from SplitMix64(seed): word = floor(next() * 1e9 / 2^64), so
every word is uniform in [0, 1e9) like the benchmark operands; the top word is
redrawn until nonzero; operand B continues the same stream after operand A.
Product lengths for the batched workload are floor(2^(4 + 10 u)) clamped to
[16, 16384] with u = next() / 2^64.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
MASK32 = np.uint64(0xFFFFFFFF)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_block(self, n):
        """The next n outputs as a uint64 array."""
        with np.errstate(over="ignore"):
            steps = (np.arange(1, n + 1, dtype=np.uint64) * GOLDEN)
            z = self.state + steps
            self.state = self.state + np.uint64(n) * GOLDEN
            return _mix(z)

    def next(self):
        return int(self.next_block(1)[0])


def to_words_1e9(z):
    """floor(z * 1e9 / 2^64) for a uint64 array, via 32-bit halves."""
    e9 = np.uint64(1000000000)
    hi = z >> np.uint64(32)
    lo = z & MASK32
    with np.errstate(over="ignore"):
        t = hi * e9 + ((lo * e9) >> np.uint64(32))
    return (t >> np.uint64(32)).astype(np.uint32)


def operand(rng, n):
    """n words uniform in [0, 1e9), top word nonzero."""
    w = to_words_1e9(rng.next_block(n))
    while n and w[-1] == 0:
        w[-1] = to_words_1e9(rng.next_block(1))[0]
    return w


def operands(seed, al, bl):
    rng = SplitMix64(seed)
    a = operand(rng, al)
    b = operand(rng, bl)
    return a, b


def batch_lengths(seed, count, lo=16, hi=16384):
    """Log-uniform operand lengths in [lo, hi] (two per product)."""
    rng = SplitMix64(seed)
    u = rng.next_block(2 * count).astype(np.float64) / 2.0**64
    lens = np.floor(2.0 ** (4 + 10 * u)).astype(np.int64)
    return np.clip(lens, lo, hi).reshape(count, 2)


def batch_operands(seed, count, lo=16, hi=16384):
    """The batched workload: per product log-uniform lengths (al, bl) in
    [lo, hi] as batch_lengths, then the words of A and B of each product in
    turn from the same SplitMix64 stream (after all lengths)."""
    rng = SplitMix64(seed)
    u = rng.next_block(2 * count).astype(np.float64) / 2.0**64
    lens = np.clip(np.floor(2.0 ** (4 + 10 * u)).astype(np.int64), lo, hi).reshape(count, 2)
    pairs = []
    for al, bl in lens:
        pairs.append((operand(rng, int(al)), operand(rng, int(bl))))
    return pairs


def words_to_int(w):
    return int.from_bytes(np.ascontiguousarray(w, dtype="<u4").tobytes(), "little")


def int_to_words(x, n):
    return np.frombuffer(int(x).to_bytes(4 * n, "little"), dtype="<u4").copy()


# Modular fingerprints of word arrays for size-independent product checks:
# (A mod p)(B mod p) == C mod p for Mersenne primes p = 2^61-1 and 2^31-1,
# where 2^(32 i) mod p = 2^((32 i) mod e) makes the residue a vectorised sum.
def _residue_mersenne(w, e):
    p = (1 << e) - 1
    w = np.ascontiguousarray(w, dtype=np.uint32).astype(np.uint64)
    k = ((np.arange(w.size, dtype=np.uint64) * np.uint64(32)) % np.uint64(e)).astype(np.uint64)
    mask = np.uint64(p)
    with np.errstate(over="ignore"):
        lo = (w << k) & mask
        sh = np.where(k + np.uint64(32) > np.uint64(e), np.uint64(e) - np.minimum(k, np.uint64(e)), np.uint64(63))
        hi = np.where(k + np.uint64(32) > np.uint64(e), w >> sh, np.uint64(0))
    t = lo + hi  # each < 2^e + 2^32, reduce once
    t = np.where(t >= mask, t - mask, t)
    s_hi = int((t >> np.uint64(32)).sum(dtype=np.uint64))
    s_lo = int((t & MASK32).sum(dtype=np.uint64))
    return ((s_hi << 32) + s_lo) % p


def product_fingerprint_ok(a, b, c):
    """True when c == a*b modulo 2^61-1 and 2^31-1 (c holds all len(a)+len(b) words)."""
    for e in (61, 31):
        p = (1 << e) - 1
        if (_residue_mersenne(a, e) * _residue_mersenne(b, e)) % p != _residue_mersenne(c, e):
            return False
    return True
