"""e pipeline host pieces (CPU): the series length of main.c:38-61."""
import math

import bluecrunch_amd as bc


def test_e_terms_known_sizes():
    # series lengths for 1e8 and 1e9 digits (p = digits / 9 + 1 words)
    assert bc.e_terms(11111112) == 15821668
    assert bc.e_terms(111111112) == 138854691


def test_e_terms_definition():
    # smallest b with lgamma(b) >= p ln 2^32 + 1, plus 2
    for p in [1, 2, 3, 10, 100, 1000, 12345]:
        need = p * math.log(2.0**32) + 1
        b = 1
        while math.lgamma(b) < need:
            b += 1
        assert bc.e_terms(p) == b + 2, p
