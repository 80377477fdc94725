"""bluecrunch_amd -- MI355X-native FFT bigfloat multiply behind bluecrunch's C API.

The product is libbluecrunch.so (built in-tree by bluecrunch_amd/build.py): HIP
kernels for gfx950 behind the C ABI of include/bluecrunch_gpu.h, plus the
reference C API (bigfloat_mul, _fft_mul, fft.h) in C on top of it. This module
is a thin ctypes view of that library for tests, the benchmark and Python
callers. It never falls back to a CPU path: a missing or unloadable library is
an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BLUECRUNCH_LIB selects another build of the library (e.g. a variant with
# different compile-time options); default: the in-tree build.
LIB_PATH = os.environ.get("BLUECRUNCH_LIB") or os.path.join(HERE, "libbluecrunch.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p

# Every extern "C" symbol include/*.h declares, grouped by header.
EXPORTS = {
    "bluecrunch_gpu.h": [
        "bcg_init", "bcg_shutdown", "bcg_last_error", "bcg_stream", "bcg_sync", "bcg_malloc", "bcg_free",
        "bcg_memcpy_h2d", "bcg_memcpy_d2h", "bcg_memcpy_d2d", "bcg_mul_dev", "bcg_mul_host", "bcg_mul_dev_bits", "bcg_mul_dev_noalloc", "bcg_basecase_dev",
        "bcg_mul_batch_dev", "bcg_muladd_dev", "bcg_muladd_batch_dev", "bcg_batch_create", "bcg_batch_run", "bcg_batch_destroy", "bcg_fft_forward_dev", "bcg_fft_inverse_dev", "bcg_fft_pointwise_dev",
        "bcg_int_to_fft_dev", "bcg_fft_to_int_dev", "bcg_fft_forward_host", "bcg_fft_inverse_host",
        "bcg_fft_pointwise_host", "bcg_int_to_fft_host", "bcg_fft_to_int_host", "bcg_basecase_host",
        "bcg_tft_forward_host", "bcg_tft_inverse_host", "bcg_plan_query", "bcg_plan_candidates",
        "bcg_plan_force", "bcg_plan_current", "bcg_plan_digit_bias", "bcg_plan_chunk", "bcg_autotune_classes", "bcg_plan_force_batch", "bcg_plan_current_batch", "bcg_autotune", "bcg_e_terms", "bcg_e_bsr_dev", "bcg_div_fraction_dev", "bcg_radix_dev", "bcg_e_digits", "bcg_e_digits_ex", "bcg_test_fail_alloc", "bcg_bf_set", "bcg_bf_get", "bcg_bf_free", "bcg_bf_add", "bcg_bf_sub", "bcg_bf_mul", "bcg_bf_rcp", "bcg_bf_div", "bcg_prof_enable", "bcg_prof_select", "bcg_prof_read",
        "bcg_prof_reset", "bcg_stats_reset", "bcg_stats_get", "bcg_defer_begin", "bcg_defer_end",
    ],
    "fft.h": [
        "twiddle_table", "twiddle_table_size", "fft_ensure_table", "table_select", "fft_free_table",
        "fft_length", "fft_forward", "fft_inverse", "fft_pointwise", "int_to_fft", "fft_to_int",
        "tft_forward", "tft_inverse", "baileys_forward", "_baileys_forward",
    ],
    "mul.h": ["_basecase_mul", "_karatzuba_mul", "_fft_mul"],
    "bigfloat.h": [
        "bigfloat_zero", "bigfloat_iszero", "bigfloat_set", "bigfloat_free", "bigfloat_new", "bigfloat_copy",
        "bigfloat_alloc", "bigfloat_realloc", "bigfloat_mul", "bigfloat_mului", "bigfloat_mulu",
        "bigfloat_add", "bigfloat_sub", "bigfloat_div", "bigfloat_rcp",
    ],
}

PROF_KINDS = {
    "fwd_pack": 0, "fwd": 1, "mid": 2, "inv": 3, "inv_unpack": 4, "carry": 5, "other": 6, "middle": 7,
}


class MulInfo(ctypes.Structure):
    _fields_ = [("bits", ctypes.c_int), ("passes", ctypes.c_int), ("points", _sz), ("margin", ctypes.c_double),
                ("retries", ctypes.c_int), ("square", ctypes.c_int), ("group_digits", ctypes.c_int),
                ("group_stride", ctypes.c_int), ("chunks", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Bigfloat(ctypes.Structure):
    """include/bigfloat.h bigfloat_t layout."""
    _fields_ = [("coef", _u32p), ("exp", ctypes.c_int64), ("len", _sz), ("sign", ctypes.c_int)]


class DevBigfloat(ctypes.Structure):
    """bcg_bigfloat: the bigfloat_t layout with the words in device memory."""
    _fields_ = [("coef", _vp), ("exp", ctypes.c_int64), ("len", _sz), ("sign", ctypes.c_int)]


class GpuError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libbluecrunch.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpuError(f"{LIB_PATH} is missing: run `python -m bluecrunch_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    i = ctypes.c_int
    L.bcg_init.argtypes = [i]
    L.bcg_last_error.restype = ctypes.c_char_p
    L.bcg_stream.restype = _vp
    L.bcg_malloc.argtypes = [ctypes.POINTER(_vp), _sz]
    L.bcg_free.argtypes = [_vp]
    L.bcg_memcpy_h2d.argtypes = [_vp, _vp, _sz]
    L.bcg_memcpy_d2h.argtypes = [_vp, _vp, _sz]
    L.bcg_memcpy_d2d.argtypes = [_vp, _vp, _sz]
    mul_args = [_vp, _sz, _vp, _sz, _vp, _sz]
    L.bcg_mul_dev.argtypes = mul_args + [_vp, ctypes.POINTER(MulInfo)]
    L.bcg_mul_dev_bits.argtypes = mul_args + [i, _vp, ctypes.POINTER(MulInfo)]
    L.bcg_mul_dev_noalloc.argtypes = mul_args + [_vp, ctypes.POINTER(MulInfo)]
    L.bcg_mul_host.argtypes = mul_args + [ctypes.POINTER(MulInfo)]
    L.bcg_basecase_dev.argtypes = mul_args + [_vp]
    L.bcg_basecase_host.argtypes = mul_args
    L.bcg_mul_batch_dev.argtypes = [_vp, _vp, _vp] + [ctypes.POINTER(_sz)] * 6 + [_sz, _vp, _f64p]
    L.bcg_muladd_dev.argtypes = mul_args + [_vp, _sz, _vp, ctypes.POINTER(MulInfo)]
    L.bcg_muladd_batch_dev.argtypes = [_vp, _vp, _vp, _vp] + [ctypes.POINTER(_sz)] * 8 + [_sz, _vp, _f64p]
    L.bcg_batch_create.argtypes = [ctypes.POINTER(_vp), _vp, _vp, _vp, _vp] + [ctypes.POINTER(_sz)] * 8 + [_sz]
    L.bcg_batch_run.argtypes = [_vp, _vp, _f64p]
    L.bcg_batch_destroy.argtypes = [_vp]
    L.bcg_fft_forward_dev.argtypes = [_vp, _sz, _vp]
    L.bcg_fft_inverse_dev.argtypes = [_vp, _sz, _vp]
    L.bcg_fft_pointwise_dev.argtypes = [_vp, _vp, _sz, _vp]
    L.bcg_int_to_fft_dev.argtypes = [_vp, _sz, _vp, _sz, i, ctypes.POINTER(_sz), _vp]
    L.bcg_fft_to_int_dev.argtypes = [_vp, _sz, _vp, _sz, i, _f64p, _vp]
    L.bcg_fft_forward_host.argtypes = [_f64p, _sz]
    L.bcg_fft_inverse_host.argtypes = [_f64p, _sz]
    L.bcg_fft_pointwise_host.argtypes = [_f64p, _f64p, _sz]
    L.bcg_int_to_fft_host.argtypes = [_f64p, _sz, _u32p, _sz, i, ctypes.POINTER(_sz)]
    L.bcg_fft_to_int_host.argtypes = [_f64p, _sz, _u32p, _sz, i, _f64p]
    L.bcg_tft_forward_host.argtypes = [_f64p, _sz, i]
    L.bcg_tft_inverse_host.argtypes = [_f64p, _sz, i]
    L.bcg_plan_query.argtypes = [_sz, _sz, ctypes.POINTER(i), ctypes.POINTER(_sz), ctypes.POINTER(i)]
    L.bcg_plan_candidates.argtypes = [_sz, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
    L.bcg_plan_force.argtypes = [_sz, i]
    L.bcg_plan_digit_bias.argtypes = [i]
    L.bcg_plan_chunk.argtypes = [_sz, i]
    L.bcg_autotune_classes.argtypes = [_sz, _sz, i, ctypes.POINTER(i)]
    L.bcg_plan_force_batch.argtypes = [_sz, i]
    L.bcg_plan_current_batch.argtypes = [_sz]
    L.bcg_plan_current.argtypes = [_sz]
    L.bcg_autotune.argtypes = [_sz, _sz, i, i, ctypes.POINTER(i), _f64p, i]
    L.bcg_e_terms.argtypes = [_sz]
    L.bcg_e_terms.restype = _sz
    L.bcg_e_bsr_dev.argtypes = [_sz, ctypes.POINTER(_vp), ctypes.POINTER(_sz), ctypes.POINTER(_sz)]
    L.bcg_div_fraction_dev.argtypes = [_vp, _sz, _vp, _sz, _sz, ctypes.POINTER(_vp)]
    L.bcg_radix_dev.argtypes = [_vp, _sz, _sz, _vp]
    L.bcg_e_digits.argtypes = [_sz, ctypes.c_char_p, _sz, _f64p]
    L.bcg_e_digits_ex.argtypes = [_sz, ctypes.c_char_p, _sz, _f64p, i]
    L.bcg_test_fail_alloc.argtypes = [ctypes.c_long]
    L.bcg_prof_enable.argtypes = [i]
    L.bcg_prof_select.argtypes = [ctypes.c_uint]
    L.bcg_prof_read.argtypes = [i, ctypes.POINTER(ctypes.c_long), _f64p]
    L.bcg_stats_get.argtypes = [ctypes.POINTER(Stats)]
    L.bcg_defer_begin.argtypes = []
    L.bcg_defer_end.argtypes = [ctypes.POINTER(ctypes.c_double)]
    # reference C API
    L.fft_length.argtypes = [_sz]; L.fft_length.restype = _sz
    L.table_select.argtypes = [_sz]; L.table_select.restype = _sz
    L.fft_ensure_table.argtypes = [_sz]; L.fft_ensure_table.restype = None
    L.fft_forward.argtypes = [_f64p, _sz]; L.fft_forward.restype = None
    L.fft_inverse.argtypes = [_f64p, _sz]; L.fft_inverse.restype = None
    L.baileys_forward.argtypes = [_f64p, _sz]; L.baileys_forward.restype = None
    L.fft_pointwise.argtypes = [_f64p, _f64p, _sz]; L.fft_pointwise.restype = None
    L.int_to_fft.argtypes = [_f64p, _sz, _u32p, _sz, i]; L.int_to_fft.restype = _sz
    L.fft_to_int.argtypes = [_f64p, _sz, _u32p, _sz, i]; L.fft_to_int.restype = None
    L.tft_forward.argtypes = [_f64p, _sz, i]; L.tft_forward.restype = None
    L.tft_inverse.argtypes = [_f64p, _sz, i]; L.tft_inverse.restype = None
    for name in ("_basecase_mul", "_karatzuba_mul"):
        getattr(L, name).argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz]
        getattr(L, name).restype = None
    L._fft_mul.argtypes = [_u32p, _sz, _u32p, _sz, _u32p, _sz]
    L._fft_mul.restype = i
    for name in ("bigfloat_mul", "bigfloat_add", "bigfloat_sub", "bigfloat_div"):
        getattr(L, name).argtypes = [ctypes.POINTER(Bigfloat)] * 3 + [_sz]
        getattr(L, name).restype = None
    L.bigfloat_rcp.argtypes = [ctypes.POINTER(Bigfloat)] * 2 + [_sz]
    L.bigfloat_rcp.restype = None
    L.bcg_bf_set.argtypes = [ctypes.POINTER(DevBigfloat), _vp, _sz, ctypes.c_int64, i]
    L.bcg_bf_get.argtypes = [ctypes.POINTER(DevBigfloat), _vp]
    L.bcg_bf_free.argtypes = [ctypes.POINTER(DevBigfloat)]
    for name in ("bcg_bf_add", "bcg_bf_sub", "bcg_bf_mul", "bcg_bf_div"):
        getattr(L, name).argtypes = [ctypes.POINTER(DevBigfloat)] * 3 + [_sz]
    L.bcg_bf_rcp.argtypes = [ctypes.POINTER(DevBigfloat)] * 2 + [_sz]
    L.bigfloat_free.argtypes = [ctypes.POINTER(Bigfloat)]
    L.bigfloat_free.restype = None
    _lib = L
    return L


def check(rc, what="bcg call"):
    if rc != 0:
        msg = lib().bcg_last_error()
        raise GpuError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def exported_symbols():
    """Names every header declares, all of which the library must export."""
    return [s for names in EXPORTS.values() for s in names]


def _u32(a):
    return a.ctypes.data_as(_u32p)


def _f64(a):
    return a.ctypes.data_as(_f64p)


# ------------------------------------------------------------ host-array API

def mul(a, b, cl=None, square=None):
    """Low cl words of a*b on the GPU (host arrays in, host array out)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    sq = square if square is not None else (b is a)
    b = a if sq else np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size if cl is None else cl
    c = np.zeros(max(cl, 1), dtype=np.uint32)
    info = MulInfo()
    check(lib().bcg_mul_host(c.ctypes.data, cl, a.ctypes.data, a.size, b.ctypes.data, b.size,
                             ctypes.byref(info)), "bcg_mul_host")
    return c[:cl], info.as_dict()


def basecase(a, b, cl=None):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size if cl is None else cl
    c = np.zeros(max(cl, 1), dtype=np.uint32)
    check(lib().bcg_basecase_host(c.ctypes.data, cl, a.ctypes.data, a.size, b.ctypes.data, b.size),
          "bcg_basecase_host")
    return c[:cl]


def plan(al, bl):
    bits, pts, passes = ctypes.c_int(), _sz(), ctypes.c_int()
    check(lib().bcg_plan_query(al, bl, ctypes.byref(bits), ctypes.byref(pts), ctypes.byref(passes)), "plan")
    return {"bits": bits.value, "points": pts.value, "passes": passes.value}


# Reference C API (include/fft.h, include/mul.h, include/bigfloat.h)

def fft_length(n):
    return lib().fft_length(n)


def fft_forward(z):
    assert z.dtype == np.complex128 and z.flags.c_contiguous
    lib().fft_forward(_f64(z.view(np.float64)), z.size)
    return z


def fft_inverse(z):
    assert z.dtype == np.complex128 and z.flags.c_contiguous
    lib().fft_inverse(_f64(z.view(np.float64)), z.size)
    return z


def baileys_forward(z):
    lib().baileys_forward(_f64(z.view(np.float64)), z.size)
    return z


def tft_forward(z, length, k):
    """fft.h tft_forward on a complex128 array of 2^k points (in place)."""
    assert z.dtype == np.complex128 and z.size >= (1 << k)
    lib().tft_forward(_f64(z.view(np.float64)), length, k)
    return z


def tft_inverse(z, length, k):
    assert z.dtype == np.complex128 and z.size >= (1 << k)
    lib().tft_inverse(_f64(z.view(np.float64)), length, k)
    return z


def fft_pointwise(a, b):
    lib().fft_pointwise(_f64(a.view(np.float64)), _f64(b.view(np.float64)), a.size)
    return a


def int_to_fft(length, words, bits):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    pts = (32 * words.size + bits - 1) // bits if words.size else 0
    z = np.zeros(max(length, pts, 1), dtype=np.complex128)
    n = lib().int_to_fft(_f64(z.view(np.float64)), length, _u32(words), words.size, bits)
    return z, n


def fft_to_int(z, length, nwords, bits):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(max(nwords, 1), dtype=np.uint32)
    lib().fft_to_int(_f64(z.view(np.float64)), length, _u32(out), nwords, bits)
    return out[:nwords]


def fft_to_int_margin(z, length, nwords, bits):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(max(nwords, 1), dtype=np.uint32)
    m = ctypes.c_double()
    check(lib().bcg_fft_to_int_host(_f64(z.view(np.float64)), length, _u32(out), nwords, bits, ctypes.byref(m)),
          "fft_to_int")
    return out[:nwords], m.value


def fft_mul(a, b, cl=None):
    """The reference _fft_mul entry point; returns (words, bits)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    cl = a.size + b.size if cl is None else cl
    c = np.zeros(max(cl, 1), dtype=np.uint32)
    bits = lib()._fft_mul(_u32(c), cl, _u32(a), a.size, _u32(b), b.size)
    return c[:cl], bits


def _bigfloat_host_call(fn, a, b, p, inplace=None):
    """A reference bigfloat operation of the library (host bigfloat_t) on
    (words, exp, sign) triples; b None for the one-operand rcp, `b is a` passes
    the same bigfloat twice. inplace="a" makes the target the bigfloat a."""
    L = lib()
    libc = ctypes.CDLL(None)
    libc.malloc.restype = _vp
    libc.malloc.argtypes = [_sz]

    def mk(t):
        w, e, s = t
        w = np.ascontiguousarray(w, dtype=np.uint32)
        bf = Bigfloat()
        if w.size:
            p_ = libc.malloc(4 * w.size)
            ctypes.memmove(p_, w.ctypes.data, 4 * w.size)
            bf.coef = ctypes.cast(p_, _u32p)
        else:
            bf.coef = ctypes.cast(None, _u32p)
        bf.exp, bf.len, bf.sign = e, w.size, s
        return bf

    A = mk(a)
    B = None if b is None else (A if b is a else mk(b))
    if inplace == "a":
        T = A
    else:
        T = Bigfloat()
        T.coef = ctypes.cast(None, _u32p)
        T.len, T.exp, T.sign = 0, 0, 1
    if B is None:
        fn(ctypes.byref(T), ctypes.byref(A), p)
    else:
        fn(ctypes.byref(T), ctypes.byref(A), ctypes.byref(B), p)
    words = np.ctypeslib.as_array(T.coef, shape=(T.len,)).copy() if T.len else np.zeros(0, np.uint32)
    res = (words, T.exp, T.sign)
    for bf in {id(x): x for x in (A, B, T) if x is not None}.values():
        L.bigfloat_free(ctypes.byref(bf))
    return res


def bigfloat_mul(a, b, p=0, inplace=None):
    """Reference bigfloat_mul on (words, exp, sign) triples; `a is b` is squaring.
    inplace="a" makes the target the same bigfloat as a (test/test_regression.c)."""
    return _bigfloat_host_call(lib().bigfloat_mul, a, b, p, inplace)


def bigfloat_add(a, b, p=0, inplace=None):
    return _bigfloat_host_call(lib().bigfloat_add, a, b, p, inplace)


def bigfloat_sub(a, b, p=0, inplace=None):
    return _bigfloat_host_call(lib().bigfloat_sub, a, b, p, inplace)


def bigfloat_div(a, b, p=0, inplace=None):
    return _bigfloat_host_call(lib().bigfloat_div, a, b, p, inplace)


def bigfloat_rcp(a, p, inplace=None):
    return _bigfloat_host_call(lib().bigfloat_rcp, a, None, p, inplace)


class DeviceBigfloat:
    """A bcg_bigfloat (device words) made from a (words, exp, sign) triple."""

    def __init__(self, t=None):
        self.bf = DevBigfloat()
        self.bf.sign = 1
        if t is not None:
            w, e, s = t
            w = np.ascontiguousarray(w, dtype=np.uint32)
            check(lib().bcg_bf_set(ctypes.byref(self.bf), w.ctypes.data if w.size else None, w.size, e, s),
                  "bcg_bf_set")

    def triple(self):
        n = self.bf.len
        w = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().bcg_bf_get(ctypes.byref(self.bf), w.ctypes.data), "bcg_bf_get")
        return w[:n], self.bf.exp, self.bf.sign

    def free(self):
        if lib() is not None and self.bf.coef:
            lib().bcg_bf_free(ctypes.byref(self.bf))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bf_op(name, t, a, b=None, p=0):
    """t = a (op) b on device bigfloats (DeviceBigfloat objects; t may be a or b)."""
    fn = getattr(lib(), "bcg_bf_" + name)
    if b is None:
        check(fn(ctypes.byref(t.bf), ctypes.byref(a.bf), p), "bcg_bf_" + name)
    else:
        check(fn(ctypes.byref(t.bf), ctypes.byref(a.bf), ctypes.byref(b.bf), p), "bcg_bf_" + name)
    return t


# ---------------------------------------------------------- device buffers

class DeviceWords:
    """Device array of uint32 words allocated through the library."""

    def __init__(self, n):
        self.n = int(n)
        p = _vp()
        check(lib().bcg_malloc(ctypes.byref(p), max(4 * self.n, 4)), "bcg_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.uint32)
        d = cls(arr.size)
        if arr.size:
            check(lib().bcg_memcpy_h2d(d.ptr, arr.ctypes.data, 4 * arr.size), "h2d")
        return d

    def zero(self):
        """Clear the words (from a host array held for the whole copy)."""
        z = np.zeros(max(self.n, 1), dtype=np.uint32)
        if self.n:
            check(lib().bcg_memcpy_h2d(self.ptr, z.ctypes.data, 4 * self.n), "h2d")

    def to_host(self, n=None):
        n = self.n if n is None else n
        out = np.zeros(max(n, 1), dtype=np.uint32)
        if n:
            check(lib().bcg_memcpy_d2h(out.ctypes.data, self.ptr, 4 * n), "d2h")
        return out[:n]

    def free(self):
        if self.ptr:
            lib().bcg_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mul_dev(c, cl, a, al, b, bl, bits=0, stream=None):
    """Device-pointer multiply (ints are raw device addresses)."""
    info = MulInfo()
    check(lib().bcg_mul_dev_bits(c, cl, a, al, b, bl, bits, stream, ctypes.byref(info)), "bcg_mul_dev")
    return info.as_dict()


def plan_candidates(n):
    """Pass splits available for an n-point transform: list of (passes, L, C)."""
    L = lib()
    count = L.bcg_plan_candidates(n, -1, None, None, None)
    if count < 0:
        check(count, "plan_candidates")
    out = []
    for i in range(count):
        P = ctypes.c_int()
        Ls = (ctypes.c_int * 3)()
        Cs = (ctypes.c_int * 3)()
        L.bcg_plan_candidates(n, i, ctypes.byref(P), Ls, Cs)
        out.append((P.value, list(Ls)[:P.value], list(Cs)[:P.value]))
    return out


def plan_force(n, index):
    check(lib().bcg_plan_force(n, index), "plan_force")


def defer_begin():
    """Open a deferral scope: this thread's products stop reading their
    margins back one by one (bcg_defer_begin)."""
    check(lib().bcg_defer_begin(), "defer_begin")


def defer_end():
    """Close the scope; returns its worst rounding margin (results of the
    scope are exact only below 0.25)."""
    w = ctypes.c_double()
    check(lib().bcg_defer_end(ctypes.byref(w)), "defer_end")
    return w.value


def plan_digit_bias(extra_bits):
    """Plan digit sizes up to extra_bits above the rounding bound (0: normal);
    exercises the rounding-margin retry."""
    check(lib().bcg_plan_digit_bias(extra_bits), "plan_digit_bias")


def plan_chunk(nbytes, streams=2):
    """Chunked middle passes of three-pass products: group size in bytes of
    operand points (0: whole-transform passes) and 1 to 4 streams."""
    check(lib().bcg_plan_chunk(int(nbytes), int(streams)), "plan_chunk")


def plan_current(n):
    """Index into plan_candidates(n) of the split n-point transforms use now."""
    r = lib().bcg_plan_current(n)
    check(min(r, 0), "plan_current")
    return r


def autotune_classes(lo=64, hi=1 << 26, reps=2):
    """Tune the pass split of batched products for every length class in
    [lo, hi]; returns {N: (passes, L, C)} of the classes with a choice."""
    n = ctypes.c_int()
    check(lib().bcg_autotune_classes(lo, hi, reps, ctypes.byref(n)), "autotune_classes")
    out = {}
    b = 32
    while 2 * b <= hi:
        for N in (2 * b, 3 * b):
            if lo <= N <= hi and len(plan_candidates(N)) > 1:
                out[N] = plan_candidates(N)[plan_current_batch(N)]
        b *= 2
    return out


def plan_force_batch(n, index):
    check(lib().bcg_plan_force_batch(n, index), "plan_force_batch")


def plan_current_batch(n):
    r = lib().bcg_plan_current_batch(n)
    check(min(r, 0), "plan_current_batch")
    return r


def autotune(al, bl, square=False, reps=2):
    """Time every pass split for an al x bl product; returns (chosen, times_ms)."""
    n = len(plan_candidates(plan(al, bl)["points"]))
    times = (ctypes.c_double * max(n, 1))()
    chosen = ctypes.c_int()
    check(lib().bcg_autotune(al, bl, 1 if square else 0, reps, ctypes.byref(chosen), times, n), "autotune")
    return chosen.value, list(times)[:n]


def prof_enable(on=True):
    check(lib().bcg_prof_enable(1 if on else 0), "prof_enable")


def prof_select(kinds=None):
    """Time only these kinds (names or numbers) while profiling; None: all."""
    mask = 0xFFFFFFFF
    if kinds is not None:
        mask = 0
        for k in kinds:
            mask |= 1 << (PROF_KINDS[k] if isinstance(k, str) else k)
    check(lib().bcg_prof_select(mask), "prof_select")


def prof_read(kind):
    n, ms = ctypes.c_long(), ctypes.c_double()
    check(lib().bcg_prof_read(PROF_KINDS[kind] if isinstance(kind, str) else kind, ctypes.byref(n),
                              ctypes.byref(ms)), "prof_read")
    return n.value, ms.value


def prof_reset():
    check(lib().bcg_prof_reset(), "prof_reset")


class Stats(ctypes.Structure):
    """bcg_stats: running statistics over FFT products since the last reset."""
    _fields_ = [("worst_margin", ctypes.c_double), ("products", ctypes.c_size_t), ("max_points", ctypes.c_size_t),
                ("bits_at_max", ctypes.c_int), ("margin_at_max", ctypes.c_double), ("retries", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def stats_reset():
    check(lib().bcg_stats_reset(), "stats_reset")


def stats():
    """Worst margin, product count and the digit size / margin at the longest
    transform since the last stats_reset()."""
    st = Stats()
    check(lib().bcg_stats_get(ctypes.byref(st)), "stats_get")
    return st.as_dict()


def sync():
    check(lib().bcg_sync(), "sync")


def mul_batch(pairs, cls=None, adds=None):
    """Batched products a*b (+ d) of host word arrays in one call
    (bcg_muladd_batch_dev); returns (list of product arrays, worst margin)."""
    count = len(pairs)
    al = np.array([len(a) for a, _ in pairs], dtype=np.uint64)
    bl = np.array([len(b) for _, b in pairs], dtype=np.uint64)
    cl = (al + bl) if cls is None else np.array(cls, dtype=np.uint64)
    offs = lambda n: np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64) if len(n) else n  # noqa: E731
    a_off, b_off, c_off = offs(al), offs(bl), offs(cl)
    cat = lambda xs: np.concatenate([np.ascontiguousarray(x, np.uint32) for x in xs]) if xs else np.zeros(0, np.uint32)  # noqa: E731
    dA, dB = DeviceWords.from_host(cat([a for a, _ in pairs])), DeviceWords.from_host(cat([b for _, b in pairs]))
    dC = DeviceWords(int(cl.sum()))
    if adds is not None:
        dl = np.array([len(d) for d in adds], dtype=np.uint64)
        d_off = offs(dl)
        dD = DeviceWords.from_host(cat(list(adds)))
    keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in (c_off, cl, a_off, al, b_off, bl)]
    if adds is not None:
        keep += [np.ascontiguousarray(d_off, np.uint64), np.ascontiguousarray(dl, np.uint64)]
    ptrs = [k.ctypes.data_as(ctypes.POINTER(_sz)) for k in keep]
    if adds is None:
        ptrs += [None, None]
    m = ctypes.c_double()
    check(lib().bcg_muladd_batch_dev(dC.ptr, dA.ptr, dB.ptr, dD.ptr if adds is not None else None, *ptrs,
                                     count, None, ctypes.byref(m)), "bcg_muladd_batch_dev")
    C = dC.to_host()
    out = [C[int(o):int(o) + int(n)].copy() for o, n in zip(c_off, cl)]
    return out, m.value


class Batch:
    """A batch of products on device words captured once into a hipGraph
    (bcg_batch_create) and replayed with run(); offsets/lengths in words."""

    def __init__(self, dC, dA, dB, c_off, cl, a_off, al, b_off, bl, dD=None, d_off=None, dl=None):
        self._keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in (c_off, cl, a_off, al, b_off, bl)]
        if dD is not None:
            self._keep += [np.ascontiguousarray(d_off, np.uint64), np.ascontiguousarray(dl, np.uint64)]
        ptrs = [k.ctypes.data_as(ctypes.POINTER(_sz)) for k in self._keep]
        if dD is None:
            ptrs += [None, None]
        h = _vp()
        check(lib().bcg_batch_create(ctypes.byref(h), dC.ptr, dA.ptr, dB.ptr, dD.ptr if dD is not None else None,
                                     *ptrs, len(self._keep[0])), "bcg_batch_create")
        self.handle = h.value

    def run(self):
        """Replay the graph; returns the worst rounding margin."""
        m = ctypes.c_double()
        check(lib().bcg_batch_run(self.handle, None, ctypes.byref(m)), "bcg_batch_run")
        return m.value

    def close(self):
        if self.handle:
            check(lib().bcg_batch_destroy(self.handle), "bcg_batch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def muladd(a, b, d, cl=None):
    """a*b + d (low cl words) on the GPU through bcg_muladd_dev."""
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    d = np.ascontiguousarray(d, np.uint32)
    cl = a.size + b.size + 1 if cl is None else cl
    dA, dB, dD, dC = DeviceWords.from_host(a), DeviceWords.from_host(b), DeviceWords.from_host(d), DeviceWords(cl)
    info = MulInfo()
    check(lib().bcg_muladd_dev(dC.ptr, cl, dA.ptr, a.size, dB.ptr, b.size, dD.ptr, d.size, None, ctypes.byref(info)),
          "bcg_muladd_dev")
    return dC.to_host(), info.as_dict()


def e_terms(words):
    """Series terms for `words` words of precision (main.c:38-61)."""
    return lib().bcg_e_terms(words)


def e_bsr(terms):
    """(P, Q) of sum_{k=1}^{terms} 1/k! = P/Q by binary splitting on the GPU."""
    p = _vp()
    pl, ql = _sz(), _sz()
    check(lib().bcg_e_bsr_dev(terms, ctypes.byref(p), ctypes.byref(pl), ctypes.byref(ql)), "bcg_e_bsr_dev")
    n = pl.value + ql.value
    out = np.zeros(max(n, 1), dtype=np.uint32)
    check(lib().bcg_memcpy_d2h(out.ctypes.data, p.value, 4 * n), "d2h")
    check(lib().bcg_free(p.value), "free")
    return out[:pl.value], out[pl.value:n]


def div_fraction(p_words, q_words, nw):
    """floor(P * 2^(32 nw) / Q) (nw words, exact) on the GPU, for P < Q."""
    dP, dQ = DeviceWords.from_host(p_words), DeviceWords.from_host(q_words)
    f = _vp()
    check(lib().bcg_div_fraction_dev(dP.ptr, len(p_words), dQ.ptr, len(q_words), nw, ctypes.byref(f)),
          "bcg_div_fraction_dev")
    out = np.zeros(nw, dtype=np.uint32)
    check(lib().bcg_memcpy_d2h(out.ctypes.data, f.value, 4 * nw), "d2h")
    check(lib().bcg_free(f.value), "free")
    return out


def radix(fraction_words, groups):
    """Base-1e9 digits of a binary fraction (host words, value F / 2^(32 n))."""
    dF = DeviceWords.from_host(fraction_words)
    dD = DeviceWords(groups)
    check(lib().bcg_radix_dev(dF.ptr, len(fraction_words), groups, dD.ptr), "bcg_radix_dev")
    return dD.to_host()


def e_digits(digits, per_multiply=False, raw=False):
    """("2." + digits decimal digits of e, timings dict) computed on the GPU.

    per_multiply: binary-splitting products one call each instead of batched.
    raw: return a read-only memoryview of the bytes (no zero fill, no copy,
    no decode) for very long outputs."""
    buf = np.empty(digits + 3, dtype=np.uint8)
    ms = (ctypes.c_double * 4)()
    check(lib().bcg_e_digits_ex(digits, buf.ctypes.data_as(ctypes.c_char_p), digits + 3, ms,
                                1 if per_multiply else 0), "bcg_e_digits")
    times = {"bsr_ms": ms[0], "div_ms": ms[1], "radix_ms": ms[2], "total_ms": ms[3]}
    # the text ends at its terminator: digits + 2 characters, fewer for a few
    # small digit counts (as the reference's e.txt)
    n = digits + 2 if digits >= 100 else buf[:digits + 3].tobytes().index(b"\0")
    if raw:
        return memoryview(buf)[:n].toreadonly(), times
    return buf[:n].tobytes().decode(), times
