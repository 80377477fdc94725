"""The length-class autotuner (bcg_autotune_classes): every pass split it may
pick for a batched length class, the two-pass 8192-point split beside the
one-workgroup kernel included, gives exact products; after tuning, batches,
captured batches and the e pipeline stay exact."""
import hashlib

import numpy as np
import pytest

from helpers import exact_product
from bluecrunch_amd.workload import SplitMix64, batch_operands, operand

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("words", [3500, 4000, 4500, 9000, 40000])
def test_every_batched_split_is_exact(bc, words):
    n = bc.plan(words, words)["points"]
    cands = bc.plan_candidates(n)
    rng = SplitMix64(words)
    pairs = [(operand(rng, words - 3 * k), operand(rng, words - 5 * k)) for k in range(8)]
    pairs.append((operand(rng, 300), operand(rng, 200)))  # another class in the same call
    want = [exact_product(a, b) for a, b in pairs]
    try:
        for i in range(len(cands)):
            bc.plan_force_batch(n, i)
            out, worst = bc.mul_batch(pairs)
            bad = [k for k, (c, w) in enumerate(zip(out, want)) if not np.array_equal(c, w)]
            assert not bad, (i, cands[i], bad)
            assert worst < 0.25
    finally:
        bc.plan_force_batch(n, -1)


def test_single_product_two_pass_8192(bc):
    # the two-pass split of 8192 points for one product (bcg_plan_force)
    a, b = SplitMix64(3), None
    x, y = operand(a, 4000), operand(a, 3999)
    n = bc.plan(4000, 3999)["points"]
    assert n == 8192
    cands = bc.plan_candidates(n)
    assert cands[0][0] == 1 and any(c[0] == 2 for c in cands)
    try:
        for i in range(len(cands)):
            bc.plan_force(n, i)
            c, info = bc.mul(x, y)
            assert np.array_equal(c, exact_product(x, y)) and info["passes"] == cands[i][0]
    finally:
        bc.plan_force(n, -1)


def test_autotune_classes_then_workloads_exact(bc, golden):
    table = bc.autotune_classes(64, 1 << 18, reps=1)
    assert 8192 in table and len(table) >= 8
    for n, (passes, L, C) in table.items():
        assert passes in (1, 2, 3) and int(np.prod(L)) == n
    pairs = batch_operands(7, 256)
    out, worst = bc.mul_batch(pairs)
    bad = [k for k, ((x, y), c) in enumerate(zip(pairs, out)) if not np.array_equal(c, exact_product(x, y))]
    assert not bad and worst < 0.25
    txt, _ = bc.e_digits(10 ** 6)
    assert hashlib.md5(txt.encode()).hexdigest() == golden["e_md5"]["1000000"]
