#!/usr/bin/env python3
"""Writes tests/golden/vectors.json.

Two kinds of fixtures, both data:
  * known-answer vectors the reference's own tests assert (values copied from
    the assertions of test/test_fft.c, test/test_fft_to_int.c,
    test/test_int_to_fft.c, test/test_tft.c, test/test_regression.c and
    test/test_bigfloat_exp.c), and the e digit MD5s of md5sums.txt;
  * exact products of seeded operands computed here with Python integers
    (arithmetic, not the reference), stored as SHA-256 digests of the
    little-endian product words so large cases stay small.
Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from bluecrunch_amd.workload import operands, words_to_int, int_to_words  # noqa: E402


def sha(words):
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()


def main():
    v = {}
    # test/test_fft.c:24-38 -- fft_forward({1,2,3,4}) and fft_inverse of that
    v["fft4"] = {"input": [[1, 0], [2, 0], [3, 0], [4, 0]],
                 "forward": [[10, 0], [-2, 0], [-2, -2], [-2, 2]],
                 "inverse": [[4, 0], [8, 0], [12, 0], [16, 0]]}
    # test/test_fft_to_int.c:20-94 -- int_to_fft known answers (length arg 4)
    v["int_to_fft"] = [
        {"words": [0x1234abcd], "bits": 12, "points": [0xbcd, 0x34a, 0x12]},
        {"words": [0x1234abcd, 0xdeadbeef], "bits": 12, "points": [0xbcd, 0x34a, 0xf12, 0xbee, 0xead, 0xd]},
        {"words": [0x1234abcd, 0xdeadbeef, 0xcafebabe], "bits": 12,
         "points": [0xbcd, 0x34a, 0xf12, 0xbee, 0xead, 0xbed, 0xeba, 0xcaf]},
        {"words": [0x1234abcd], "bits": 8, "points": [0xcd, 0xab, 0x34, 0x12]},
    ]
    # test/test_int_to_fft.c:8-56 -- fft_to_int with points scaled by 16, length 16
    v["fft_to_int"] = [
        {"points": [0xbcd, 0x34a, 0x12], "length": 16, "bits": 12, "words": [0x1234abcd]},
        {"points": [0xbcd, 0x34a, 0xa12, 0xbba, 0xdcc, 0xd], "length": 16, "bits": 12,
         "words": [0x1234abcd, 0xddccbbaa]},
        {"points": [0xbcd, 0x34a, 0xa12, 0xbba, 0xdcc, 0x11d, 0x322, 0x443], "length": 16, "bits": 12,
         "words": [0x1234abcd, 0xddccbbaa, 0x44332211]},
    ]
    # test/test_conversion.c:18-45 -- round trip driven by glibc rand()
    v["conversion"] = {"seed": 17699823, "runs": 10000, "words": 64, "buffer": 300}
    # test/test_tft.c -- forward: (input, len, k, expected); inverse: (input, len, k, expected)
    v["tft_forward"] = [
        {"input": [[1, 0], [2, 0], [3, 0], [0, 0]], "len": 3, "k": 2,
         "expected": [[6, 0], [2, 0], [-2, 2], [0, 0]]},
        {"input": [[1, 0], [2, 0], [3, 0], [4, 0], [5, 0], [0, 0], [0, 0], [0, 0]], "len": 5, "k": 3,
         "expected": [[15, 0], [3, 0], [3, -2], [3, 2], [-5.4142135623730949, 7.2426406871192857], [0, 0], [0, 0],
                      [0, 0]]},
    ]
    v["tft_inverse"] = [
        {"input": [[6, 0], [2, 0], [-2, 2], [0, 0]], "len": 3, "k": 2, "expected": [4, 8, 12, 0]},
        {"input": [[28, 0], [4, 0], [-4, 4], [-4, -4], [-9.6568542494923797, -4], [1.6568542494923797, -4],
                   [1.6568542494923801, 4], [0, 0]], "len": 7, "k": 3,
         "expected": [8, 16, 24, 32, 40, 48, 56, 0]},
        {"input": [[55, 0], [-7, 0], [-10, 17], [-10, -17], [-9.36396103, 20.53553391], [3.36396103, 13.46446609],
                   [0, 0], [0, 0]], "len": 6, "k": 3, "expected": [16, 88, 136, 56, 40, 104, 0, 0]},
        {"input": [[209, 0], [-49, 0], [28, -61], [28, 61], [-95.13351365, 117.21677477], [0, 0], [0, 0], [0, 0]],
         "len": 5, "k": 3, "expected": [8, 272, 208, 760, 424, 0, 0, 0]},
        {"input": [[647, 0], [-29, 0], [81, -74], [81, 74], [-42.52186130, 87.1543289], [220.52186130, -23.1543289],
                   [220.52186130, 23.1543289], [-42.52186130, -87.1543289], [-118.78486382, 197.7215088],
                   [-48.89637704, -63.4372376], [-59.61723845, 56.2080022], [67.29847932, 21.5077265],
                   [0, 0], [0, 0], [0, 0], [0, 0]], "len": 12, "k": 4,
         "expected": [x * 16 for x in [51, 34, 26, 95, 53, 93, 41, 37, 91, 5, 47, 74]] + [0, 0, 0, 0]},
        {"input": [[521, 0], [3, 0], [128, -5], [128, 5], [6.26850660, -15.70710678], [171.73149340, -14.29289322],
                   [171.73149340, 14.29289322], [6.26850660, 15.70710678], [-52.61287345, 301.23603015]]
         + [[0, 0]] * 7, "len": 9, "k": 4,
         "expected": [x * 16 for x in [51, 34, 26, 95, 53, 93, 41, 37, 91]] + [0] * 7},
    ]
    # test/test_regression.c:40-67 -- in-place squaring a = a*a
    v["regression_square"] = {"a": [0x44210000, 0xb1514c28, 0xce88deb8, 0x00000002],
                              "expected": [0x00000000, 0xf4718c41, 0x21c8adaa, 0x28b6543c, 0x68a5465b, 0xe0c40a81,
                                           0x7]}
    # test/test_bigfloat_exp.c:27-111 -- 0x1234^0xd7 by square-and-multiply
    v["bigfloat_exp"] = {"base": 0x1234, "exp": 0xd7, "len": 82,
                         "hex": "10298d253fecb89d491933590a085405544b42f1d9c940ca666a0b296a5c661fbcac60e9bbb478e26b3e"
                                "640d61c2832cd6b61f7d67ea02df7c14ce8279e958b612d26d54ae6e5a7ce6e7c4b24156ff511648e672"
                                "df968070750ec664af4ed6b56a11935c75d52781051282e03e2b55c4e9f9eef659e327a3ad9e7c171134"
                                "5d7e26c9db89d3eb94e59f53f2c8924737736cbd1321f8d80ea626aaa0a8b9691847816ebfa322982ebd"
                                "48bca0475d049403e81e57362d491b6625b25de3b364fe1ee0c4ddf80121e042a81231e258d627343b35"
                                "c08b2e51eaa632b034342209cb69c7171de9c9bdc47e90d39de0b73703f46d15c8cc1bf1914405dc7de7"
                                "08e4c8814342828b0ba1f04c4cc854a672040e4a93c9400000000" + "0" * 99,
                         "asserted_top": [0x10298d25, 0x3fecb89d, 0x49193359, 0x0a085405, 0x544b42f1, 0xd9c940ca,
                                          0x666a0b29, 0x6a5c661f, 0xbcac60e9, 0xbbb478e2, 0x6b3e640d]}
    # md5sums.txt:1-8 -- MD5 of "2." + N decimal digits of e
    v["e_md5"] = {
        "1000": "b5fb3d734b44d720dbf6acc49124d958",
        "10000": "99e76ec81c90890bda140f53f687ed0e",
        "100000": "c93ccb588383340b30a5e78b0368ce09",
        "1000000": "94cf5c8f7d082cc826b11219d4e0d1cb",
        "10000000": "b7d9c0783489a8aa5cfdeec5b4f18a25",
        "100000000": "01e38fd5f4975672f00944e901b4679d",
        "200000000": "851809b31a46989bcb3b5e2812345ed8",
        "1000000000": "644034cb947736dc0a1f7b3b56c28c72",
    }
    # exact products of seeded operands (SplitMix64, words in [0,1e9))
    prods = []
    cases = [(11, 1, 1), (12, 3, 5), (13, 40, 59), (14, 50, 50), (15, 99, 1), (16, 100, 100), (17, 257, 1000),
             (18, 1024, 1024), (19, 3000, 5000), (20, 4096, 4096), (21, 8191, 8193), (22, 16384, 16384),
             (23, 1, 16384), (24, 12345, 777), (25, 65536, 65536)]
    for seed, al, bl in cases:
        a, b = operands(seed, al, bl)
        c = words_to_int(a) * words_to_int(b)
        prods.append({"seed": seed, "al": al, "bl": bl, "sha256": sha(int_to_words(c, al + bl))})
        sq = words_to_int(a) ** 2
        prods[-1]["square_sha256"] = sha(int_to_words(sq, 2 * al))
    v["products"] = prods
    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(v, f, indent=1)
    print("wrote", os.path.join(HERE, "vectors.json"))


if __name__ == "__main__":
    main()
