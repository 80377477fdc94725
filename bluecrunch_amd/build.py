"""Build libbluecrunch.so in-tree: hipcc for the gfx950 kernels and the C ABI,
gcc for the C host API, linked into one shared library next to this file.

Run as `python -m bluecrunch_amd.build` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libbluecrunch.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BLUECRUNCH_ARCH", "gfx950")

# Packed FP32 instructions are switched off for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCLUDE}"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build(verbose=True, jobs=None, out=None, build_dir=None, defines=()):
    """Compile every csrc/*.hip unit (in parallel) and api.c, link libbluecrunch.so
    (or `out`); `defines` are extra -D options for the HIP units."""
    from concurrent.futures import ThreadPoolExecutor

    bdir = build_dir or BUILD_DIR
    lib_out = out or LIB
    os.makedirs(bdir, exist_ok=True)
    units = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(bdir, u[:-4] + ".o") for u in units]
    dflags = [f"-D{d}" for d in defines]
    cmds = [[HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, u), "-o", o] for u, o in zip(units, objs)]
    api_o = os.path.join(bdir, "api.o")
    cmds.append(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "api.c"), "-o", api_o])
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    os.makedirs(os.path.dirname(lib_out), exist_ok=True)
    tmp = lib_out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, api_o, "-o", tmp, "-lm", "-pthread"])
    os.replace(tmp, lib_out)
    return lib_out


if __name__ == "__main__":
    build()
    sys.exit(0)
