"""Binary splitting of e on the GPU against exact integers
(docs/binary-splitting.txt example, closed forms P = sum (k+1)...T, Q = T!)."""
import math

import numpy as np
import pytest

from helpers import words_to_int

pytestmark = pytest.mark.gpu


def exact_pq(T):
    q = math.factorial(T)
    p, run = 0, 1
    for k in range(T, 0, -1):  # P = sum_{k=1}^{T} (k+1)...T
        p += run
        run *= k
    return p, q


def test_bsr_docs_example(bc):
    # docs/binary-splitting.txt:44-92: 5 terms -> P = 206, Q = 120
    P, Q = bc.e_bsr(5)
    assert words_to_int(P) == 206 and words_to_int(Q) == 120


@pytest.mark.parametrize("T", [1, 2, 15, 16, 17, 33, 100, 1000, 4097, 30000, 200000])
def test_bsr_exact(bc, T):
    P, Q = bc.e_bsr(T)
    p, q = exact_pq(T)
    assert words_to_int(Q) == q
    assert words_to_int(P) == p
    assert P[-1] != 0 and Q[-1] != 0


@pytest.mark.parametrize("T,nw", [(5, 2), (50, 4), (1000, 40), (30000, 900), (200000, 6000)])
def test_fraction_division(bc, T, nw):
    P, Q = bc.e_bsr(T)
    p, q = words_to_int(P), words_to_int(Q)
    F = bc.div_fraction(P, Q, nw)
    want = (p << (32 * nw)) // q
    assert abs(words_to_int(F) - want) <= 8
