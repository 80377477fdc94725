"""The drop-in recipe (the reference builds without its multiply sources and
links libbluecrunch.so), checked on the reference's own callers:
every C file the recipe keeps (main.c, bench.c, bigfloat/*.c but mul.c,
util/bench.c, fft/fftnew.c and the test programs) compiles with the reference's
warning flags and its own headers first, and every program links against
libbluecrunch.so with no unresolved symbol; the hot-path symbols resolve to the
library, not to a reference object. Nothing built here is run.

Compile flags: the reference Makefile's CFLAGS (Makefile:8) without the three
parts this image cannot honour: -fcilkplus (no Cilk Plus runtime: the
keywords come from the serial-elision stand-in in tests/refbuild/cilk),
-fprofile-use=. (no profile data) and -flto (irrelevant to linking). Skipped
where the reference tree is not present."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import bluecrunch_amd as bc

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
CFLAGS = ["-std=c11", "-Wall", "-Werror", "-O3", "-no-pie", "-mavx", "-funroll-loops"]
# the files the recipe drops from the reference build (replaced by the
# library): the multiply path, and optionally also sum, reciprocal and quotient
DROPPED = {"fft/fft.c", "fft/baileys.c", "fft/tft.c", "fft/pointwise.c", "fft/convert.c", "fft/table.c",
           "bigfloat/mul.c"}
DROPPED_ARITH = DROPPED | {"bigfloat/addsub.c", "bigfloat/rcp.c", "bigfloat/div.c"}
ARITH = {"bigfloat_add", "bigfloat_sub", "bigfloat_rcp", "bigfloat_div"}
# symbols of the multiply path every program must take from the library
HOT = {"bigfloat_mul", "_fft_mul", "_karatzuba_mul", "_basecase_mul", "fft_forward", "fft_inverse",
       "fft_pointwise", "int_to_fft", "fft_to_int", "fft_ensure_table", "fft_free_table", "fft_length",
       "tft_forward", "tft_inverse", "baileys_forward", "_baileys_forward", "table_select", "bigfloat_mului",
       "bigfloat_mulu", "twiddle_table", "twiddle_table_size"}

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "main.c")) or not shutil.which("gcc"),
                                reason="reference tree or gcc not present")


def _kept_sources(dropped=DROPPED):
    out = []
    for d in ("bigfloat", "fft", "util"):
        for f in sorted(os.listdir(os.path.join(REF, d))):
            rel = f"{d}/{f}"
            if f.endswith(".c") and rel not in dropped:
                out.append(rel)
    return out


@pytest.fixture(scope="module")
def objects(tmp_path_factory):
    if not os.path.exists(bc.LIB_PATH):
        pytest.fail("libbluecrunch.so is not built")
    d = tmp_path_factory.mktemp("reflink")
    inc = [f"-I{os.path.join(HERE, 'refbuild')}", f"-I{REF}/include", f"-I{REF}"]
    progs = ["main.c", "bench.c"] + [f"test/{f}" for f in sorted(os.listdir(os.path.join(REF, "test")))
                                     if f.endswith(".c")]
    srcs = _kept_sources() + progs

    def cc(rel):
        obj = d / (rel.replace("/", "_")[:-2] + ".o")
        r = subprocess.run(["gcc", *CFLAGS, *inc, "-c", os.path.join(REF, rel), "-o", str(obj)],
                           capture_output=True, text=True)
        return rel, obj, r

    with ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(cc, srcs))
    errors = {rel: r.stderr for rel, _, r in res if r.returncode}
    assert not errors, errors
    objs = {rel: obj for rel, obj, _ in res}
    return d, objs, {p: objs[p] for p in progs}


def _link(d, prog_obj, lib_objs, name):
    libdir = os.path.dirname(bc.LIB_PATH)
    exe = d / name
    r = subprocess.run(["gcc", *CFLAGS, "-o", str(exe), str(prog_obj), *map(str, lib_objs), f"-L{libdir}",
                        "-lbluecrunch", f"-Wl,-rpath,{libdir}", "-lm"], capture_output=True, text=True)
    return exe, r


def _undefined(exe):
    out = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1].split("@")[0] for line in out.splitlines() if line.strip()}


def _defined(obj):
    out = subprocess.run(["nm", "--defined-only", str(obj)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("variant", ["multiply", "arith"])
def test_kept_objects_define_no_replaced_symbol(objects, variant):
    _, objs, _ = objects
    dropped, hot = (DROPPED, HOT) if variant == "multiply" else (DROPPED_ARITH, HOT | ARITH)
    for rel in _kept_sources(dropped):
        assert not (_defined(objs[rel]) & hot), (rel, _defined(objs[rel]) & hot)


@pytest.mark.parametrize("prog", ["main.c", "bench.c", "test/test_regression.c", "test/test_bigfloat_exp.c",
                                  "test/test_fft.c", "test/test_fft_to_int.c", "test/test_int_to_fft.c",
                                  "test/test_conversion.c", "test/test_tft.c", "test/test_baileys.c",
                                  "test/test_fft_benchmark.c", "test/test_radix_conv.c",
                                  "test/test_bigfloat_log.c", "test/test_test.c"])
@pytest.mark.parametrize("variant", ["multiply", "arith"])
def test_program_links_against_the_library(objects, prog, variant):
    d, objs, progs = objects
    dropped = DROPPED if variant == "multiply" else DROPPED_ARITH
    lib_objs = [objs[r] for r in _kept_sources(dropped)]
    exe, r = _link(d, progs[prog], lib_objs, prog.replace("/", "_")[:-2] + "_" + variant)
    assert r.returncode == 0, r.stderr
    und = _undefined(exe)
    provided = set(bc.exported_symbols())
    # what the program takes from the library is what the library exports
    assert (und & (HOT | ARITH)) <= provided
    if variant == "arith" and prog == "main.c":  # e's division and sums from the library too
        assert {"bigfloat_div", "bigfloat_add"} <= und
    if prog == "main.c":  # e_BSR, rcp, div, exp and radix all multiply through the library
        assert {"bigfloat_mul", "fft_ensure_table", "fft_free_table", "bigfloat_mului", "bigfloat_mulu"} <= und
    if prog == "bench.c":
        assert {"_fft_mul", "_karatzuba_mul", "_basecase_mul", "fft_forward", "fft_inverse"} <= und
