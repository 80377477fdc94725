"""ctypes view of the oracle library (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference FFT multiply,
used by tests/ as the parity checker and by bench.py as the CPU baseline. The
product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefcpu.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


class RefBigfloat(ctypes.Structure):
    """Layout of the reference bigfloat_t (include/bigfloat.h:6-13)."""
    _fields_ = [("coef", _u32p), ("exp", ctypes.c_int64), ("len", ctypes.c_size_t), ("sign", ctypes.c_int)]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sz = ctypes.c_size_t
        L.ref_fft_length.argtypes = [sz]; L.ref_fft_length.restype = sz
        L.ref_table_select_raw.argtypes = [sz]; L.ref_table_select_raw.restype = sz
        L.ref_ensure_table.argtypes = [sz]; L.ref_ensure_table.restype = None
        L.ref_table_size.restype = ctypes.c_int
        L.ref_fft_forward.argtypes = [_f64p, sz]; L.ref_fft_forward.restype = None
        L.ref_fft_inverse.argtypes = [_f64p, sz]; L.ref_fft_inverse.restype = None
        L.ref_fft_pointwise.argtypes = [_f64p, _f64p, sz]; L.ref_fft_pointwise.restype = None
        L.ref_int_to_fft.argtypes = [_f64p, sz, _u32p, sz, ctypes.c_int]; L.ref_int_to_fft.restype = sz
        L.ref_fft_to_int.argtypes = [_f64p, sz, _u32p, sz, ctypes.c_int]; L.ref_fft_to_int.restype = None
        L.ref_fft_to_int_margin.argtypes = [_f64p, sz, _u32p, sz, ctypes.c_int]
        L.ref_fft_to_int_margin.restype = ctypes.c_double
        L.ref_sched_bits.argtypes = [sz]; L.ref_sched_bits.restype = ctypes.c_int
        for name in ("ref_basecase_mul", "ref_karatzuba_mul"):
            f = getattr(L, name)
            f.argtypes = [_u32p, sz, _u32p, sz, _u32p, sz]
            f.restype = None
        L.ref_fft_mul_bits.argtypes = [_u32p, sz, _u32p, sz, _u32p, sz, ctypes.c_int, _f64p]
        L.ref_fft_mul_bits.restype = ctypes.c_int
        L.ref_fft_mul.argtypes = [_u32p, sz, _u32p, sz, _u32p, sz]; L.ref_fft_mul.restype = ctypes.c_int
        for name in ("ref_bigfloat_mul", "ref_bigfloat_add", "ref_bigfloat_sub", "ref_bigfloat_div"):
            getattr(L, name).argtypes = [ctypes.POINTER(RefBigfloat)] * 3 + [sz]
            getattr(L, name).restype = None
        L.ref_bigfloat_rcp.argtypes = [ctypes.POINTER(RefBigfloat)] * 2 + [sz]
        L.ref_bigfloat_rcp.restype = None
        L.ref_bigfloat_ucmp.argtypes = [ctypes.POINTER(RefBigfloat)] * 2
        L.ref_bigfloat_ucmp.restype = ctypes.c_int
        L.ref_e_terms.argtypes = [sz]
        L.ref_e_terms.restype = sz
        L.ref_e_digits.argtypes = [sz, ctypes.c_char_p, sz, _f64p]
        L.ref_e_digits.restype = sz
        L.ref_num_threads.restype = ctypes.c_int
        L.ref_set_threads.argtypes = [ctypes.c_int]
        L.ref_set_threads.restype = None
        _lib = L
    return _lib


def _u32(a):
    return a.ctypes.data_as(_u32p)


def _f64(a):
    return a.ctypes.data_as(_f64p)


def fft_length(n):
    return lib().ref_fft_length(n)


def fft_forward(z):
    """In-place forward transform of a complex128 array (reference order/sign)."""
    assert z.dtype == np.complex128 and z.flags.c_contiguous
    lib().ref_fft_forward(_f64(z.view(np.float64)), z.size)
    return z


def fft_inverse(z):
    assert z.dtype == np.complex128 and z.flags.c_contiguous
    lib().ref_fft_inverse(_f64(z.view(np.float64)), z.size)
    return z


def fft_pointwise(a, b):
    lib().ref_fft_pointwise(_f64(a.view(np.float64)), _f64(b.view(np.float64)), a.size)
    return a


def int_to_fft(length, words, bits, capacity=None):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    al = words.size
    pts = (32 * al + bits - 1) // bits if al else 0
    cap = max(length, pts) if capacity is None else capacity
    z = np.zeros(cap, dtype=np.complex128)
    n = lib().ref_int_to_fft(_f64(z.view(np.float64)), length, _u32(words), al, bits)
    return z, n


def fft_to_int(z, length, nwords, bits):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(max(nwords, 1), dtype=np.uint32)
    m = lib().ref_fft_to_int_margin(_f64(z.view(np.float64)), length, _u32(out), nwords, bits)
    return out[:nwords], m


def basecase_mul(a, b, cl=None):
    a = np.ascontiguousarray(a, dtype=np.uint32); b = np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size if cl is None else cl
    c = np.zeros(max(cl, 1), dtype=np.uint32)
    lib().ref_basecase_mul(_u32(c), cl, _u32(a), a.size, _u32(b), b.size)
    return c[:cl]


def karatzuba_mul(a, b, cl=None):
    a = np.ascontiguousarray(a, dtype=np.uint32); b = np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size if cl is None else cl
    c = np.zeros(max(cl, 1), dtype=np.uint32)
    lib().ref_karatzuba_mul(_u32(c), cl, _u32(a), a.size, _u32(b), b.size)
    return c[:cl]


def fft_mul(a, b, bits=0, square=False):
    """Reference _fft_mul on word arrays; returns (words, bits, margin)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = a if square else np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size
    c = np.zeros(cl, dtype=np.uint32)
    m = ctypes.c_double(0.0)
    used = lib().ref_fft_mul_bits(_u32(c), cl, _u32(a), a.size, _u32(b), b.size, bits, ctypes.byref(m))
    return c, used, m.value


def _bigfloat_call(fn, a, b, p):
    """Run a reference bigfloat operation on (words, exp, sign) triples (b None
    for the one-operand rcp); returns the target as a triple."""
    keep = []

    def mk(t):
        w, e, s = t
        w = np.ascontiguousarray(w, dtype=np.uint32)
        keep.append(w)
        bf = RefBigfloat()
        bf.coef = _u32(w) if w.size else ctypes.cast(None, _u32p)
        bf.exp, bf.len, bf.sign = e, w.size, s
        return bf

    A = mk(a)
    B = None if b is None else (A if b is a else mk(b))
    T = RefBigfloat()
    T.coef = ctypes.cast(None, _u32p)
    T.len, T.exp, T.sign = 0, 0, 1
    if B is None:
        fn(ctypes.byref(T), ctypes.byref(A), p)
    else:
        fn(ctypes.byref(T), ctypes.byref(A), ctypes.byref(B), p)
    words = np.ctypeslib.as_array(T.coef, shape=(T.len,)).copy() if T.len else np.zeros(0, np.uint32)
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    if T.coef:
        libc.free(ctypes.cast(T.coef, ctypes.c_void_p))
    return words, T.exp, T.sign


def bigfloat_mul(a, b, p=0):
    """Reference bigfloat_mul on (words, exp, sign) triples; returns the same."""
    return _bigfloat_call(lib().ref_bigfloat_mul, a, b, p)


def bigfloat_add(a, b, p=0):
    return _bigfloat_call(lib().ref_bigfloat_add, a, b, p)


def bigfloat_sub(a, b, p=0):
    return _bigfloat_call(lib().ref_bigfloat_sub, a, b, p)


def bigfloat_rcp(a, p):
    return _bigfloat_call(lib().ref_bigfloat_rcp, a, None, p)


def bigfloat_div(a, b, p=0):
    return _bigfloat_call(lib().ref_bigfloat_div, a, b, p)


def e_digits(digits):
    """The reference's e.txt for `digits` decimals ("2." + digits), computed by
    the restated e pipeline (oracle/refe.c); returns (bytes, timings dict)."""
    buf = ctypes.create_string_buffer(digits + 32)
    ms = (ctypes.c_double * 4)()
    n = lib().ref_e_digits(digits, buf, digits + 32, ms)
    return buf.raw[:n], {"bsr_ms": ms[0], "div_ms": ms[1], "radix_ms": ms[2], "total_ms": ms[3]}


def e_terms(p):
    return lib().ref_e_terms(p)


def num_threads():
    return lib().ref_num_threads()


def set_threads(n):
    """OpenMP threads used by later oracle calls."""
    lib().ref_set_threads(int(n))
