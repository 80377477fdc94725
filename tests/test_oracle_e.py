"""Pins the oracle's restatement of the reference's e pipeline (oracle/refe.c:
main.c:38-167 with bigfloat/div.c, addsub.c, radix.c and string.c): its e.txt
against the reference's md5sums.txt, digit for digit against e computed with
Python integers, and its term counts against the library's. CPU only."""
import hashlib

import numpy as np
import pytest

from test_gpu_e import e_digits_exact


@pytest.mark.parametrize("digits", [1000, 10000, 100000, 1000000])
def test_oracle_e_matches_reference_md5(oracle, golden, digits):
    txt, ms = oracle.e_digits(digits)
    assert len(txt) == digits + 2 and txt[:2] == b"2."
    assert hashlib.md5(txt).hexdigest() == golden["e_md5"][str(digits)]


@pytest.mark.parametrize("digits", [5, 8, 9, 29, 99, 1001, 5000])
def test_oracle_e_digits_exact(oracle, digits):
    txt, _ = oracle.e_digits(digits)
    assert txt.decode() == "2." + e_digits_exact(digits)[1:]


def test_oracle_e_short_texts(oracle):
    # bigfloat/string.c:84-112 prints at most (digits + 2) / 8 base-1e9 groups
    # for an output of digits + 2 characters (main.c:145), so for a few small
    # digit counts e.txt holds fewer digits than asked: 1 -> "2.", 10 -> nine
    # digits, 28 -> 27 digits; from 29 on it always holds them all
    exact = e_digits_exact(40)
    assert oracle.e_digits(1)[0] == b"2."
    assert oracle.e_digits(10)[0].decode() == "2." + exact[1:10]
    assert oracle.e_digits(28)[0].decode() == "2." + exact[1:28]
    short = [d for d in range(1, 400) if len(oracle.e_digits(d)[0]) != d + 2]
    assert short == [1, 2, 3, 4, 10, 11, 12, 19, 20, 28]


def test_e_terms_agree_with_the_library(oracle):
    import bluecrunch_amd as bc
    L = bc.lib()
    ps = list(range(1, 3000)) + [int(x) for x in np.random.default_rng(3).integers(1, 2 * 10 ** 8, 3000)]
    assert all(oracle.e_terms(p) == L.bcg_e_terms(p) for p in ps)
    # main.c:38-61 at the e to 10^8 and 10^9 digit configurations
    assert oracle.e_terms(11111112) == 15821668
    assert oracle.e_terms(111111112) == 138854691
