"""The workgroup kernels at the bottom of the e pipeline's trees, run on the
CPU: k_bsr_subtree (32 leaves of 16 terms folded and paired in LDS) and
k_radix_subtree (a node of up to 1024 base-1e9 digit groups split in LDS,
leaves by repeated multiplication), compiled with g++ against a thread-per-lane
stand-in for the HIP built-ins (tests/emu/emu.h). Their integer results are
checked against Python integers: P(a,b) = sum_{k=a+1}^{b} (k+1)...b and
Q(a,b) = (a+1)...b (main.c:63-109), and the digit groups floor(X 1e9^K) of
random fractions X, including fractions whose low words vanish (the round-up
rule of the scaled remainder tree)."""
import math
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "bluecrunch_amd", "csrc", "e_pipeline.inc")


def _cut(src, a, b):
    i = src.index(a)
    return src[i:src.index(b, i)]


@pytest.fixture(scope="module")
def drv(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("emu")
    src = open(SRC).read()
    bsr = _cut(src, "constexpr int BSR_G = 16;", "// Strip zero top words")
    bsr = bsr.replace(_cut(bsr, "__global__ void k_bsr_leaves", "// The bottom levels of the tree"), "")
    rad = ("constexpr int RADIX_LEAF = 32;\nconstexpr int RADIX_GUARD = 2;\n"
           + _cut(src, "// The bottom of the tree in one workgroup", "// Powers 1e9^K on the device"))
    with open(d / "kernels_cpu.h", "w") as f:
        f.write('#include "emu.h"\n' + bsr + "\n" + rad)
    exe = d / "drv"
    subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-w", f"-I{HERE}/emu", f"-I{d}",
                    os.path.join(HERE, "emu", "drv.cpp"), "-o", str(exe)], check=True)
    return str(exe), d


def _words(x):
    w = []
    while x:
        w.append(x & 0xffffffff)
        x >>= 32
    return w or [0]


def _pq(a, b):
    P, Q = 0, 1
    for k in range(a + 1, b + 1):
        P, Q = P * k + 1, Q * k
    return P, Q


@pytest.mark.parametrize("T", [1, 2, 15, 16, 17, 100, 511, 512, 513, 1535, 4000])
def test_bsr_subtree_exact(drv, T):
    exe, d = drv
    out = d / f"bsr{T}.txt"
    subprocess.run([exe, "bsr", str(T), str(out)], check=True)
    lines = open(out).read().split("\n")
    span = 16 * 32
    for i in range((T + span - 1) // span):
        P, Q = _pq(i * span, min(T, (i + 1) * span))
        for v, x in ((0, P), (1, Q)):
            f = list(map(int, lines[2 * i + v].split()))
            assert f[0] == len(_words(x)) and f[1:] == _words(x), (T, i, v)


def _radix_words(K):
    return math.ceil(K * math.log2(1e9) / 32) + 2


def _kh_set(Ks):
    s, out = set(Ks), set()
    while s:
        K = s.pop()
        if K <= 32:
            continue
        kh = (K + 1) // 2
        out.add(kh)
        s.update((kh, K - kh))
    return sorted(out)


@pytest.mark.parametrize("case", range(6))
def test_radix_subtree_exact(drv, case):
    exe, d = drv
    rng = random.Random(100 + case)
    Ks = rng.sample([1, 2, 31, 32, 33, 64, 65, 100, 257, 500, 511, 512, 513, 700, 1000, 1023, 1024], 4)
    if case == 5:
        Ks = [1024, 1023, 1024, 678]
    roots, d0 = [], 0
    for K in Ks:
        n = _radix_words(K)
        X = rng.getrandbits(32 * n)
        if case == 4 and K > 64:  # zero low half: the guard words of G vanish
            X = (X >> (32 * (n // 2))) << (32 * (n // 2))
        roots.append((K, n, d0, X))
        d0 += K
    pw = _kh_set(Ks)
    inp = d / f"rad{case}.in"
    with open(inp, "w") as f:
        f.write(f"{len(roots)}\n")
        for K, n, dd, X in roots:
            f.write(f"{K} {n} {dd} " + " ".join(str((X >> (32 * t)) & 0xffffffff) for t in range(n)) + "\n")
        f.write(f"{len(pw)}\n")
        for k in pw:
            w = _words(10 ** (9 * k))
            f.write(f"{k} {len(w)} " + " ".join(map(str, w)) + "\n")
    out = d / f"rad{case}.out"
    subprocess.run([exe, "radix", str(inp), str(out)], check=True)
    got = list(map(int, open(out).read().split()))
    for K, n, dd, X in roots:
        full = (X * 10 ** (9 * K)) >> (32 * n)
        want = [(full // 10 ** (9 * (K - 1 - i))) % 10 ** 9 for i in range(K)]
        assert got[dd:dd + K] == want, (K, next(i for i in range(K) if got[dd + i] != want[i]))
