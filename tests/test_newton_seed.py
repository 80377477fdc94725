"""The host long division that seeds the e pipeline's Newton reciprocal
(knuth_divide in bluecrunch_amd/csrc/e_pipeline.inc, Knuth's Algorithm D),
compiled with g++ on its own and checked against Python integer division:
random operands, divisors with a small top word (the normalisation shift),
all-ones words (the quotient-estimate corrections and the add-back step) and
the seed's own shape, floor(2^(32 (t + k)) / Q_t)."""
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "bluecrunch_amd", "csrc", "e_pipeline.inc")

DRIVER = r"""
#include <cstdint>
#include <cstdio>
#include <vector>
#include "knuth.h"
int main() {
  int un, vn;
  while (scanf("%d %d", &un, &vn) == 2) {
    std::vector<uint32_t> u(un), v(vn), q(un - vn + 1);
    for (auto &x : u) scanf("%u", &x);
    for (auto &x : v) scanf("%u", &x);
    knuth_divide(u.data(), un, v.data(), vn, q.data());
    for (auto x : q) printf("%u ", x);
    printf("\n");
  }
}
"""


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("knuth")
    src = open(SRC).read()
    i = src.index("void knuth_divide(")
    j = src.index("\n}\n", i) + 3
    (d / "knuth.h").write_text("#include <cstdint>\n#include <vector>\n" + src[i:j])
    (d / "drv.cpp").write_text(DRIVER)
    out = d / "drv"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", f"-I{d}", str(d / "drv.cpp"), "-o", str(out)], check=True)
    return str(out)


def _words(x, n):
    return [(x >> (32 * i)) & 0xffffffff for i in range(n)]


def _cases():
    rng = random.Random(3)
    cases = []
    for _ in range(300):
        vn = rng.randint(2, 12)
        un = vn + rng.randint(0, 12)
        kind = rng.randrange(4)
        if kind == 0:
            v = rng.getrandbits(32 * vn) | (1 << (32 * vn - 1 - rng.randrange(32)))
            u = rng.getrandbits(32 * un)
        elif kind == 1:  # all-ones words: maximal estimates, add-back
            v = (1 << (32 * vn)) - 1 - rng.getrandbits(8)
            u = (1 << (32 * un)) - 1 - rng.getrandbits(16)
        elif kind == 2:  # top divisor word 1: the largest normalisation shift
            v = (1 << (32 * (vn - 1))) | rng.getrandbits(32 * (vn - 1))
            u = rng.getrandbits(32 * un)
        else:  # the seed's shape: a power of two over the top words of Q
            v = rng.getrandbits(32 * vn) | (1 << (32 * vn - 1 - rng.randrange(32)))
            u = 1 << (32 * (un - 1))
        v = v or 1 << (32 * vn - 1)
        if v >> (32 * (vn - 1)) == 0:
            v |= 1 << (32 * (vn - 1))
        cases.append((u, un, v, vn))
    # the seed itself: 256 quotient words of 2^(32 (t + 256)) / Q_t, t = 259
    t, k = 259, 256
    qt = rng.getrandbits(32 * t) | (1 << (32 * t - 1))
    cases.append((1 << (32 * (t + k)), t + k + 1, qt, t))
    return cases


def test_knuth_divide_matches_python(exe):
    cases = _cases()
    inp = []
    for u, un, v, vn in cases:
        inp.append(f"{un} {vn} " + " ".join(map(str, _words(u, un))) + " " + " ".join(map(str, _words(v, vn))))
    res = subprocess.run([exe], input="\n".join(inp) + "\n", capture_output=True, text=True, check=True)
    lines = res.stdout.strip().splitlines()
    assert len(lines) == len(cases)
    for (u, un, v, vn), line in zip(cases, lines):
        got = [int(x) for x in line.split()]
        want = _words(u // v, un - vn + 1)
        assert got == want, (un, vn)
