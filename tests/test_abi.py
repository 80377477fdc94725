"""The C-ABI library: it builds, loads, and exports every symbol include/*.h
declares. No compute calls (this runs without a GPU)."""
import os
import re
import subprocess

import pytest

import bluecrunch_amd as bc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header):
    """Function and extern data names declared in a C header."""
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text):
        names.add(m.group(1))
    for m in re.finditer(r"extern\s+[^;(]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*(\[[^\]]*\])?\s*;", text):
        names.add(m.group(1))
    return names - {"if", "while", "for", "return", "sizeof"}


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", bc.LIB_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_library_built():
    assert os.path.exists(bc.LIB_PATH), "run python -m bluecrunch_amd.build"


@pytest.mark.parametrize("header", ["bluecrunch_gpu.h", "fft.h", "mul.h", "bigfloat.h"])
def test_every_declared_symbol_is_exported(header):
    decl = declared_symbols(header)
    assert decl, header
    missing = decl - exported()
    assert not missing, f"{header}: not exported: {sorted(missing)}"
    # the Python view lists the same names
    assert decl <= set(bc.EXPORTS[header]) | set(bc.EXPORTS["bluecrunch_gpu.h"])


def test_library_loads_and_binds():
    L = bc.lib()
    for name in bc.exported_symbols():
        assert hasattr(L, name), name


def test_plan_query_is_host_only():
    # plans are pure host arithmetic: length classes 2^k / 3*2^k, digits fit 2N
    for al, bl in [(100, 100), (2**17, 2**17), (2**24, 2**24), (1, 50000), (11111112, 11111112)]:
        p = bc.plan(al, bl)
        n, b = p["points"], p["bits"]
        m = n // 3 if n % 3 == 0 else n
        assert m & (m - 1) == 0
        need = -(-32 * al // b) + -(-32 * bl // b) - 1
        assert need <= 2 * n, (al, bl, p)
        assert 4 <= b <= 20 and p["passes"] in (1, 2, 3)


def test_plan_at_benchmark_size():
    # 2^24 x 2^24 words: 8-bit digits, 2^26 complex points (right-angle: 2^27 digits)
    p = bc.plan(1 << 24, 1 << 24)
    assert p["bits"] == 8 and p["points"] == 1 << 26


def test_default_split_keeps_strided_passes_wide():
    # the untuned split: every strided (non-innermost) pass moves >= 8 columns
    # per workgroup, the last pass is contiguous; forcing and resetting works
    for w in [1 << 16, 1 << 20, 1 << 22, 1 << 24, 3 << 21]:
        n = bc.plan(w, w)["points"]
        cands = bc.plan_candidates(n)
        cur = bc.plan_current(n)
        P, L, C = cands[cur]
        assert len(L) == P and all(c >= 8 for c in C[:-1]), (w, cands[cur])
        prod = 1
        for x in L:
            prod *= x
        assert prod == n
        other = (cur + 1) % len(cands)
        bc.plan_force(n, other)
        assert bc.plan_current(n) == other
        bc.plan_force(n, -1)
        assert bc.plan_current(n) == cur


def test_no_gpu_fails_loudly():
    # without a HIP device the product path returns an error code, never a CPU result
    import ctypes
    L = bc.lib()
    n = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        pass
    if n.value > 0:
        pytest.skip("a GPU is present")
    a = (ctypes.c_uint32 * 200)(*range(1, 201))
    c = (ctypes.c_uint32 * 400)()
    rc = L.bcg_mul_host(ctypes.addressof(c), 400, ctypes.addressof(a), 200, ctypes.addressof(a), 200, None)
    assert rc != 0
    assert list(c) == [0] * 400
