"""Concurrent callers: eight host threads multiply through the reference API
at once on disjoint buffers, as main.c:87-89 (Cilk-spawned e_BSR halves),
bigfloat/radix.c:269-271 and mul.c:148-150 do. ctypes releases the GIL for
the duration of each call, so the calls really overlap in the library; every
product must be exact."""
import threading

import numpy as np
import pytest

from helpers import exact_product, words_to_int
from bluecrunch_amd.workload import operands

pytestmark = pytest.mark.gpu


def test_eight_threads_fft_mul_and_bigfloat_mul(bc):
    errors = []
    sizes = [(50, 40), (700, 900), (5000, 3000), (20000, 20000), (60000, 1000), (131072, 131072)]

    def worker(tid):
        try:
            for rep in range(3):
                for k, (al, bl) in enumerate(sizes):
                    a, b = operands(1000 * tid + 10 * rep + k, al, bl)
                    c, bits = bc.fft_mul(a, b)
                    if not np.array_equal(c, exact_product(a, b)):
                        errors.append(("fft_mul", tid, rep, al, bl))
                    w, e, s = bc.bigfloat_mul((a, -3, 1), (b, tid, 0), 0)
                    x = words_to_int(a) * words_to_int(b)
                    top = (x.bit_length() + 31) // 32
                    if words_to_int(w) != x or e != tid - 3 or s != 0 or len(w) not in (top, al + bl):
                        errors.append(("bigfloat_mul", tid, rep, al, bl))
        except Exception as ex:  # noqa: BLE001
            errors.append(("exception", tid, repr(ex)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:5]


def test_threads_on_device_words_and_batches(bc, golden):
    # device-pointer products, batches and e digits from several threads
    import hashlib
    from helpers import exact_product as ex
    errors = []

    def worker(tid):
        try:
            a, b = operands(77 + tid, 30000 + tid, 20000)
            dA, dB, dC = bc.DeviceWords.from_host(a), bc.DeviceWords.from_host(b), bc.DeviceWords(50000 + tid)
            bc.mul_dev(dC.ptr, 50000 + tid, dA.ptr, 30000 + tid, dB.ptr, 20000)
            if not np.array_equal(dC.to_host(), ex(a, b)):
                errors.append(("mul_dev", tid))
            pairs = [operands(tid * 100 + k, 100 + 37 * k, 90 + 11 * k) for k in range(20)]
            out, _ = bc.mul_batch(pairs)
            if any(not np.array_equal(c, ex(x, y)) for (x, y), c in zip(pairs, out)):
                errors.append(("batch", tid))
            if tid % 2 == 0:
                txt, _ = bc.e_digits(10000)
                results[tid] = hashlib.md5(txt.encode()).hexdigest()
        except Exception as e:  # noqa: BLE001
            errors.append(("exception", tid, repr(e)))

    results = {}
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:5]
    assert set(results.values()) == {golden["e_md5"]["10000"]}
