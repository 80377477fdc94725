#!/usr/bin/env python3
"""Benchmark: bigfloat_mul of two 2^24-word operands on one MI355X.

A step is one device-resident product of two seeded 2^24-word operands
(words uniform in [0, 1e9), SplitMix64 seed 1) through the library's C ABI
(bcg_mul_dev: FP64 right-angle FFT multiply, rounding-margin check and carry
resolve). value = decimal digits per second with the north-star convention of
9 digits per operand word: 9 * 2^24 / (seconds per product). The squaring of
operand A is timed the same way and reported beside it.

Also reported: the dominant kernel's HBM roofline (algorithmic bytes per launch
over its HIP-event time on the launch stream) and the oracle (the CPU
restatement of the reference AVX path, OpenMP on the host cores) timed on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bigfloat_mul digits/s at 2^24 words (%HBM peak); e to 1e8 digits wall-s"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def algorithmic_bytes(kind, info, al, bl, cl):
    """Compulsory HBM bytes of each kernel kind in one product (16 B per complex point).

    The forward passes run once per operand (A and B, or A alone when squaring);
    a three-pass plan has one middle forward and one middle inverse pass. The
    last inverse pass writes either 2N 64-bit coefficients or, in group mode,
    one exact sum of G digits per group (group_stride words each). Per-launch
    figures divide these by the launches of that kind in one product.
    """
    n = info["points"]
    sq = info["square"]
    ops = 1 if sq else 2
    mids = max(info["passes"] - 2, 0)
    if info.get("group_digits"):
        carry_in = (2 * n // info["group_digits"]) * info["group_stride"] * 4
    else:
        carry_in = 16 * n
    if kind == "fwd_pack":
        return 4 * (al if sq else al + bl) + 16 * n * ops  # read the words, write the first-pass arrays
    if kind == "fwd":
        return 32 * n * ops * mids  # read and write each array
    if kind == "inv":
        return 32 * n * mids
    if kind == "mid":
        if info["passes"] == 1:
            return 4 * (al + (0 if sq else bl)) + 4 * cl
        return (32 if sq else 48) * n  # read A (and B) arrays, write the product array
    if kind == "inv_unpack":
        return 16 * n + carry_in  # read the product array, write coefficients / group sums
    if kind == "carry":
        return carry_in + 4 * cl  # read coefficients / group sums, write words
    if kind == "middle":
        # chunked middle section: the second forward, middle and second inverse
        # passes of every group (one span per product)
        return sum(algorithmic_bytes(k, info, al, bl, cl) for k in ("fwd", "mid", "inv"))
    return 0


# the library's default chunking of three-pass middle passes (bcg_plan_chunk):
# 128 MiB groups on two streams
CHUNK_BYTES = 128 << 20
CHUNK_STREAMS = 2


def read_traffic(kind, words):
    """HBM bytes per launch of a kernel kind measured with rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy kernel) for this workload, else None."""
    if words != 1 << 24:
        return None
    path = os.path.join(ROOT, "profiles", "bigfloat_mul_2p24w_traffic.txt")
    key = f"traffic_bytes_per_launch[{kind}]:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def plan_split(bc, n):
    """Pass lengths and columns per workgroup the library uses for n points."""
    passes, L, C = bc.plan_candidates(n)[bc.plan_current(n)]
    return {"L": list(L), "C": list(C)}


def e_pipeline(bc, digits, compare_digits, cpu_digits):
    """e = 2.71828... to `digits` decimals on the GPU (batched binary splitting,
    Newton division, radix conversion), wall time around the library call, MD5
    of the text against the reference's md5sums (tests/golden/vectors.json).
    At compare_digits the binary splitting also runs one call per product. At
    cpu_digits the oracle's restatement of the reference's e pipeline runs on
    the host cores beside a GPU run of the same size."""
    import hashlib

    with open(os.path.join(ROOT, "tests", "golden", "vectors.json")) as f:
        gold = json.load(f)["e_md5"]
    bc.e_digits(1000)  # warm-up (kernels, tables)
    bc.stats_reset()
    t0 = time.perf_counter()
    text, ms = bc.e_digits(digits, raw=True)
    wall = time.perf_counter() - t0
    st = bc.stats()
    md5 = hashlib.md5(text).hexdigest()
    # the same computation again in the same process: kernels loaded, tables
    # and memory pool warm (steady state of a long-running caller)
    t0 = time.perf_counter()
    text2, ms2 = bc.e_digits(digits, raw=True)
    wall2 = time.perf_counter() - t0
    warm_ok = hashlib.md5(text2).hexdigest() == md5
    del text2
    # the same with the default splits for batched classes (table restored after)
    wall_default = None
    saved = batch_table(bc)
    if saved:
        for n in saved:
            bc.plan_force_batch(n, -1)
        t0 = time.perf_counter()
        text3, _ = bc.e_digits(digits, raw=True)
        wall_default = time.perf_counter() - t0
        warm_ok = warm_ok and hashlib.md5(text3).hexdigest() == md5
        del text3
        for n, i in saved.items():
            bc.plan_force_batch(n, i)
    out = {"digits": digits, "wall_s": wall, "bsr_ms": ms["bsr_ms"], "div_ms": ms["div_ms"],
           "radix_ms": ms["radix_ms"], "wall_s_warm": wall2, "warm_equal": warm_ok, "md5": md5,
           "wall_s_warm_default_splits": wall_default,
           "md5_ok": (md5 == gold[str(digits)]) if str(digits) in gold else None,
           # every FFT product of the run: the worst rounding margin, and the
           # digit size and margin at the longest transform
           "products": st["products"], "worst_margin": st["worst_margin"], "max_points": st["max_points"],
           "bits_at_max": st["bits_at_max"], "margin_at_max": st["margin_at_max"]}
    if cpu_digits:
        out["cpu_baseline"] = e_cpu_baseline(bc, cpu_digits, gold)
    if compare_digits:
        print(f"bench: per-multiply binary splitting at {compare_digits} digits", file=sys.stderr, flush=True)
        _, mb = bc.e_digits(compare_digits)
        _, mp = bc.e_digits(compare_digits, per_multiply=True)
        out["bsr_compare"] = {"digits": compare_digits, "batched_bsr_ms": mb["bsr_ms"],
                              "per_multiply_bsr_ms": mp["bsr_ms"],
                              "speedup": mp["bsr_ms"] / mb["bsr_ms"] if mb["bsr_ms"] > 0 else None,
                              "batched_total_ms": mb["total_ms"], "per_multiply_total_ms": mp["total_ms"]}
    return out


def e_cpu_baseline(bc, digits, gold):
    """The reference's e pipeline (main.c:111-167) restated in the oracle
    (oracle/refe.c, OpenMP tasks at the reference's Cilk thresholds) on the
    host cores, after the twiddle tables main.c:185-209 builds before its timed
    phases; beside it the GPU pipeline at the same size (first run in this
    process at that size). Both texts are checked against md5sums."""
    import hashlib
    from oracle import refcpu

    L = refcpu.lib()
    p = 2 * digits // 9 + 10  # main.c:185-197
    length = 1
    while length < 3 * p:
        length *= 2
    L.ref_ensure_table(min(length, 100663296))
    print(f"bench: oracle e to {digits} digits on {refcpu.num_threads()} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    txt, ms = refcpu.e_digits(digits)
    cpu_s = time.perf_counter() - t0
    cpu_ok = hashlib.md5(txt).hexdigest() == gold.get(str(digits))
    del txt
    t0 = time.perf_counter()
    gtxt, gms = bc.e_digits(digits, raw=True)
    gpu_s = time.perf_counter() - t0
    gpu_ok = hashlib.md5(gtxt).hexdigest() == gold.get(str(digits))
    return {"digits": digits, "cpu_wall_s": cpu_s, "cpu_bsr_ms": ms["bsr_ms"], "cpu_div_ms": ms["div_ms"],
            "cpu_radix_ms": ms["radix_ms"], "cpu_md5_ok": cpu_ok, "gpu_wall_s": gpu_s, "gpu_md5_ok": gpu_ok,
            "speedup": cpu_s / gpu_s if gpu_s > 0 else None, "cores": refcpu.num_threads(),
            "host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": f"oracle restatement of main.c e() at {digits} digits (e_BSR, bigfloat_div, bigfloat_add, "
                      "bigfloat_radix, bigfloat_to_string), one run"}


def batch_table(bc, lo=64, hi=1 << 26):
    """{N: candidate index} of the batched length classes with a tuned choice."""
    out = {}
    b = 32
    while 2 * b <= hi:
        for n in (2 * b, 3 * b):
            if lo <= n <= hi and len(bc.plan_candidates(n)) > 1:
                out[n] = bc.plan_current_batch(n)
        b *= 2
    return out


def batch_leg(bc, count, seed):
    """The batched workload: count independent products with log-uniform operand
    lengths in [16, 2^14] words (seed 7), device-resident: replayed as one
    hipGraph (bcg_batch_create/run: products grouped by length class, one
    persistent launch per class, captured once), as one batched library call
    (bcg_mul_batch_dev, planned per call), and as one bcg_mul_dev call (one
    kernel set) per product."""
    import ctypes
    from bluecrunch_amd.workload import batch_operands

    pairs = batch_operands(seed, count)
    al = np.array([len(x) for x, _ in pairs], dtype=np.uint64)
    bl = np.array([len(y) for _, y in pairs], dtype=np.uint64)
    cl = al + bl
    off = lambda n: np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)  # noqa: E731
    a_off, b_off, c_off = off(al), off(bl), off(cl)
    dA = bc.DeviceWords.from_host(np.concatenate([x for x, _ in pairs]))
    dB = bc.DeviceWords.from_host(np.concatenate([y for _, y in pairs]))
    dC = bc.DeviceWords(int(cl.sum()))
    dC2 = bc.DeviceWords(int(cl.sum()))
    L = bc.lib()
    P = ctypes.POINTER(ctypes.c_size_t)
    arrs = [np.ascontiguousarray(v, np.uint64) for v in (c_off, cl, a_off, al, b_off, bl)]
    ptrs = [v.ctypes.data_as(P) for v in arrs]
    worst = ctypes.c_double()

    def batched():
        bc.check(L.bcg_mul_batch_dev(dC.ptr, dA.ptr, dB.ptr, *ptrs, count, None, ctypes.byref(worst)), "batch")

    info = bc.MulInfo()

    def per_call():
        for i in range(count):
            bc.check(L.bcg_mul_dev(dC2.ptr + 4 * int(c_off[i]), int(cl[i]), dA.ptr + 4 * int(a_off[i]), int(al[i]),
                                   dB.ptr + 4 * int(b_off[i]), int(bl[i]), None, ctypes.byref(info)), "mul")

    batched()
    bc.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        batched()
    bc.sync()
    tb = (time.perf_counter() - t0) / 3
    per_call()
    bc.sync()
    t0 = time.perf_counter()
    per_call()
    bc.sync()
    tp = time.perf_counter() - t0
    same = bool(np.array_equal(dC.to_host(), dC2.to_host()))
    # the same batch captured once into a hipGraph and replayed
    dC3 = bc.DeviceWords(int(cl.sum()))
    graph = bc.Batch(dC3, dA, dB, c_off, cl, a_off, al, b_off, bl)
    graph.run()
    bc.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        gw = graph.run()
    bc.sync()
    tg = (time.perf_counter() - t0) / 3
    same_graph = bool(np.array_equal(dC.to_host(), dC3.to_host()))
    graph.close()
    # the same replay with the default splits (batched length-class table
    # cleared for the measurement, then restored)
    tg_untuned = None
    saved = batch_table(bc)
    if saved:
        for n in saved:
            bc.plan_force_batch(n, -1)
        graph = bc.Batch(dC3, dA, dB, c_off, cl, a_off, al, b_off, bl)
        graph.run()
        bc.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            graph.run()
        bc.sync()
        tg_untuned = (time.perf_counter() - t0) / 3
        graph.close()
        for n, i in saved.items():
            bc.plan_force_batch(n, i)
    # exact spot checks against Python integers
    from bluecrunch_amd.workload import words_to_int, int_to_words
    C = dC.to_host()
    exact = True
    for i in range(0, count, max(1, count // 16)):
        x, y = pairs[i]
        want = int_to_words(words_to_int(x) * words_to_int(y), int(cl[i]))
        exact = exact and bool(np.array_equal(C[int(c_off[i]):int(c_off[i]) + int(cl[i])], want))
    for d in (dA, dB, dC, dC2, dC3):
        d.free()
    # what one replay must move and compute: every product reads its words and
    # writes its product; a multi-pass product also moves its transform arrays
    # (B_mul per product); 5 N log2 N FP64 flops per transform, three per product
    alg_bytes, flops, onepass_bytes = 0, 0.0, 0
    for x, y in zip(al, bl):
        pl = bc.plan(int(x), int(y))
        n = pl["points"]
        words = 4 * (2 * int(x) + 2 * int(y))
        if pl["passes"] == 1:
            alg_bytes += words
            onepass_bytes += words
        else:
            alg_bytes += 6 * 32 * n + 16 * n + words
        flops += 3 * 5.0 * n * np.log2(n)
    return {"products": count, "seed": seed, "total_operand_words": int(al.sum() + bl.sum()),
            "replay_alg_bytes": alg_bytes, "replay_GBps": alg_bytes / tg / 1e9,
            "replay_hbm_frac": alg_bytes / tg / HBM_PEAK, "replay_fp64_tflops": flops / tg / 1e12,
            "replay_fp64_frac": flops / tg / 78.6e12, "onepass_word_bytes": onepass_bytes,
            "graph_ms": tg * 1e3, "batched_ms": tb * 1e3, "per_call_ms": tp * 1e3,
            "graph_ms_default_splits": None if tg_untuned is None else tg_untuned * 1e3,
            "speedup_graph_vs_per_call": tp / tg, "speedup_batched_vs_per_call": tp / tb,
            "products_per_s": count / tg, "worst_margin": max(worst.value, gw),
            "batched_equals_per_call": same, "graph_equals_batched": same_graph, "exact_spot_checks": exact}


def cpu_baseline(words, seed, one_core_words, gpu_product=None):
    """Oracle _fft_mul (the CPU restatement of the reference AVX path, OpenMP) on
    the host cores: one warm-up then the median of 3 runs of a words x words
    product on all cores, plus one run of a smaller product on one core.
    gpu_product (optional): the GPU's product of the same operands, compared
    word for word with the oracle's."""
    from oracle import refcpu
    from bluecrunch_amd.workload import operands

    a, b = operands(seed, words, words)
    L = refcpu.lib()
    # twiddle tables are set up before timing, as in the reference main.c:209
    bits = max(L.ref_sched_bits(2 * words), 7)
    n = refcpu.fft_length(int(32.0 / bits * 2 * words))
    L.ref_ensure_table(n)
    cores = refcpu.num_threads()
    times = []
    used = bits
    for i in range(4):
        t0 = time.perf_counter()
        c, used, margin = refcpu.fft_mul(a, b)
        if i:
            times.append(time.perf_counter() - t0)
    dt = sorted(times)[1]
    words_equal = None if gpu_product is None else bool(np.array_equal(c, gpu_product))
    del c
    one = None
    if one_core_words:
        a1, b1 = operands(seed, one_core_words, one_core_words)
        refcpu.set_threads(1)
        refcpu.fft_mul(a1[:1024], b1[:1024])
        t0 = time.perf_counter()
        _, u1, _ = refcpu.fft_mul(a1, b1)
        d1 = time.perf_counter() - t0
        refcpu.set_threads(cores)
        one = {"value": 9.0 * one_core_words / d1, "unit": "digits/s", "cores": 1,
               "sample": f"one oracle _fft_mul of {one_core_words}x{one_core_words} words, b={u1}, {d1:.2f} s"}
    return {
        "value": 9.0 * words / dt,
        "unit": "digits/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle _fft_mul of {words}x{words} words (seed {seed}), b={used}, N={n} complex points: "
                  f"median of 3 runs after a warm-up, {dt:.2f} s each on {cores} threads",
        "seconds": dt,
        "one_core": one,
        # the host the baseline ran on: logical CPUs, CPUs this process may use,
        # and the OpenMP threads the oracle ran with (= "cores")
        "host_cpus": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "omp_threads": cores,
        "parity_words_equal": words_equal,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--words", type=int, default=1 << 24, help="words per operand")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-words", type=int, default=1 << 24, help="operand words of the CPU baseline sample")
    ap.add_argument("--cpu-one-core-words", type=int, default=1 << 20,
                    help="operand words of the one-core CPU sample (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default pass split)")
    ap.add_argument("--no-square", action="store_true", help="skip the squaring runs (profiler runs)")
    ap.add_argument("--e-digits", type=int, default=10 ** 8, help="digits of e for the e leg (0: skip)")
    ap.add_argument("--batch", type=int, default=4096, help="products in the batched leg (0: skip)")
    ap.add_argument("--e-compare-digits", type=int, default=10 ** 8,
                    help="digits of e for the batched vs per-multiply binary-splitting comparison (0: skip)")
    ap.add_argument("--e-cpu-digits", type=int, default=10 ** 7,
                    help="digits of e for the CPU baseline of the e pipeline (oracle on the host cores; 0: skip)")
    ap.add_argument("--split", default=None,
                    help="force the pass split, e.g. 64,256,4096 (skips the autotuner); for profiler runs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB three times with a plain elementwise kernel "
                         "(reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the FFT multiply runs on a single MI355X "
              "(the library drives one device per process; sharding a product over GPUs is not implemented).",
              file=sys.stderr)
        sys.exit(2)

    import bluecrunch_amd as bc
    from bluecrunch_amd.workload import operands, product_fingerprint_ok

    bc.check(bc.lib().bcg_init(0), "bcg_init")
    al = bl = args.words
    cl = al + bl
    a, b = operands(args.seed, al, bl)
    dA = bc.DeviceWords.from_host(a)
    dB = bc.DeviceWords.from_host(b)
    dC = bc.DeviceWords(cl)
    dS = bc.DeviceWords(2 * al)

    def step():
        return bc.mul_dev(dC.ptr, cl, dA.ptr, al, dB.ptr, bl)

    def step_sq():
        return bc.mul_dev(dS.ptr, 2 * al, dA.ptr, al, dA.ptr, al)

    # untimed set-up: the autotuner times every pass split of this transform
    # length once and keeps the fastest (the product and the square share N)
    tune_ms = None
    if args.split:
        want = [int(x) for x in args.split.split(",")]
        n = bc.plan(al, bl)["points"]
        idx = [i for i, c in enumerate(bc.plan_candidates(n)) if list(c[1]) == want]
        if not idx:
            sys.exit(f"no pass split {want} for {n} points")
        bc.plan_force(n, idx[0])
    elif not args.no_tune:
        t_tune = time.perf_counter()
        bc.autotune(al, bl, reps=2)
        tune_ms = (time.perf_counter() - t_tune) * 1e3
    # the length classes of batched products (the batch leg, the e pipeline's
    # tree levels): pass split per class, timed on batches of that class
    class_table, class_ms = None, None
    if not args.no_tune and (args.batch or args.e_digits):
        print("bench: length-class autotuner", file=sys.stderr, flush=True)
        t_tune = time.perf_counter()
        class_table = bc.autotune_classes(64, 1 << 26, reps=2)
        class_ms = (time.perf_counter() - t_tune) * 1e3
    for _ in range(args.warmup):
        info = step()
        if not args.no_square:
            step_sq()
    bc.sync()

    ok = None
    if not args.no_check:
        info = step()
        c = dC.to_host()
        ok = bool(product_fingerprint_ok(a, b, c))
        if not args.no_square:
            step_sq()
            ok = ok and bool(product_fingerprint_ok(a, a, dS.to_host()))

    kinds = ("fwd_pack", "fwd", "mid", "inv", "inv_unpack", "carry", "middle")

    def read_kernels(steps, inf=None):
        inf = inf or info
        out = {}
        for kind in kinds:
            n, ms = bc.prof_read(kind)
            if n:
                per = ms / n
                byt = algorithmic_bytes(kind, inf, al, bl, cl) * steps / n  # per launch
                out[kind] = {"launches": n, "avg_ms": per, "alg_bytes": byt, "GBps": byt / (per * 1e-3) / 1e9}
        return out

    # per-kernel breakdown (untimed): every kernel bracketed by HIP events
    prof_steps = min(args.steps, 5)
    bc.prof_reset()
    bc.prof_select(None)
    bc.prof_enable(True)
    for _ in range(prof_steps):
        info = step()
    bc.sync()
    bc.prof_enable(False)
    kern = read_kernels(prof_steps)
    bc.prof_reset()
    # with chunked middle passes the launches of the two streams overlap and a
    # launch's span is no measure of its kernel: the same kernels are also
    # timed over whole transforms (chunking off), one launch per pass
    kern_whole = None
    if info.get("chunks"):
        bc.plan_chunk(0)
        bc.prof_enable(True)
        for _ in range(prof_steps):
            info_whole = step()
        bc.sync()
        bc.prof_enable(False)
        kern_whole = read_kernels(prof_steps, info_whole)
        bc.prof_reset()
        bc.plan_chunk(CHUNK_BYTES, CHUNK_STREAMS)
    # the dominant unit: the chunked middle section when the product runs its
    # middle passes in groups on two streams (their launches overlap, so a
    # launch's own span is no measure of it), else the costliest kernel kind
    if "middle" in kern:
        dom = "middle"
    else:
        dom = max(kern, key=lambda k: kern[k]["avg_ms"] * kern[k]["launches"]) if kern else None

    # timed region: K products; only the dominant kernel kind carries HIP events
    # (recorded on its launch stream), so the events cost the wall time little
    bc.prof_select([dom] if dom else [])
    bc.prof_enable(True)
    bc.sync()
    t0 = time.perf_counter()
    worst = 0.0
    for _ in range(args.steps):
        info = step()
        worst = max(worst, info["margin"])
    bc.sync()
    t1 = time.perf_counter()
    bc.prof_enable(False)
    live = read_kernels(args.steps)
    bc.prof_reset()
    bc.prof_select(None)

    # squaring, timed the same way (not profiled)
    s_info, s_worst, t2, t3 = None, 0.0, 0.0, 0.0
    if not args.no_square:
        bc.sync()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            s_info = step_sq()
            s_worst = max(s_worst, s_info["margin"])
        bc.sync()
        t3 = time.perf_counter()

    sec = (t1 - t0) / args.steps
    sec_sq = (t3 - t2) / args.steps
    digits = 9.0 * al
    value = digits / sec

    roof = None
    if dom and dom in live:
        d = live[dom]
        ach = d["alg_bytes"] / (d["avg_ms"] * 1e-3)
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": read_traffic(dom, al)}
        if dom == "middle":
            roof["span"] = (f"second forward, middle and second inverse pass kernels of {info['chunks']} groups "
                            "on two streams, fork to join, one span per product")
    # FFT passes (everything but the carry) per product: their bytes over the
    # product's wall time less the carry kernels (passes may overlap in time)
    fft_bytes = sum(v["alg_bytes"] * v["launches"] for k, v in kern.items() if k not in ("carry", "middle")) / prof_steps
    carry_ms = kern["carry"]["avg_ms"] * kern["carry"]["launches"] / prof_steps if "carry" in kern else 0.0

    # host arrays in and out (PCIe copies included): never the headline value
    host_ms = None
    if not args.no_check:
        bc.mul(a, b)
        t4 = time.perf_counter()
        for _ in range(3):
            bc.mul(a, b)
        host_ms = (time.perf_counter() - t4) / 3 * 1e3

    print("bench: e leg", file=sys.stderr, flush=True)
    e_leg = e_pipeline(bc, args.e_digits, args.e_compare_digits, 0 if args.no_cpu else args.e_cpu_digits) \
        if args.e_digits else None
    print("bench: batch leg", file=sys.stderr, flush=True)
    b_leg = batch_leg(bc, args.batch, 7) if args.batch else None

    cpu = None
    if not args.no_cpu:
        # dC holds the last timed product of the same seeded operands
        gpu_c = dC.to_host() if (args.cpu_words == al and not args.no_check) else None
        cpu = cpu_baseline(args.cpu_words, args.seed, args.cpu_one_core_words, gpu_c)
        del gpu_c

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "digits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"bigfloat_mul {al}x{bl} words, words uniform in [0,1e9) from SplitMix64 seed {args.seed}, "
                        "device-resident operands; plus squaring of operand A",
            "words": al,
            "digits_per_product": digits,
            "bits_per_digit": info["bits"],
            "points": info["points"],
            "passes": info["passes"],
            "worst_margin": worst,
            "pass_split": plan_split(bc, info["points"]),
            "autotune_ms": tune_ms,
            "autotune_classes_ms": class_ms,
            # batched length class -> pass lengths / columns per workgroup the tuner kept
            "autotune_classes": None if class_table is None else {
                str(n): "x".join(map(str, L)) + "/" + "x".join(map(str, C[:-1] or [1]))
                for n, (passes, L, C) in sorted(class_table.items())},
        },
        "square": None if s_info is None else {
            "ms_per_step": sec_sq * 1e3, "digits_per_s": digits / sec_sq, "worst_margin": s_worst,
            "bits_per_digit": s_info["bits"], "points": s_info["points"]},
        "host_pointer": None if host_ms is None else {
            "ms_per_product": host_ms, "digits_per_s": digits / (host_ms * 1e-3),
            "note": "bcg_mul_host: operands copied in and the product copied out over PCIe (8 MiB chunks through pinned staging; the product into a fresh array each call)"},
        "fft_passes_hbm_frac": (fft_bytes / (sec - carry_ms * 1e-3)) / HBM_PEAK if sec > carry_ms * 1e-3 else None,
        "kernels": kern,
        # chunked products: the per-launch spans above overlap across the two
        # streams; these are the same kernels over whole transforms
        "kernels_whole_passes": kern_whole,
        "roofline": roof,
        "cpu_baseline": cpu,
        "e": e_leg,
        "batch": b_leg,
        "speedup_vs_cpu_baseline": (value / cpu["value"]) if cpu else None,
        "parity_fingerprint_ok": ok,
        # the GPU product against the oracle's product, word for word (None
        # when the CPU leg is skipped or runs another size)
        "parity_words_equal": None if cpu is None else cpu["parity_words_equal"],
    }
    print(json.dumps(out), flush=True)

    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)


def copy_calibration(mib):
    """Three launches of the library's plain vector copy kernel (bcg_memcpy_d2d)
    of `mib` MiB: each reads and writes exactly mib MiB, the yardstick that
    scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import bluecrunch_amd as bc

    words = mib * (1 << 20) // 4
    src = bc.DeviceWords(words)
    dst = bc.DeviceWords(words)
    for _ in range(3):
        bc.check(bc.lib().bcg_memcpy_d2d(dst.ptr, src.ptr, words * 4), "memcpy_d2d")
    print(f"copy calibration: 3 x {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
