"""ctypes access to libkukan_hip.so, the MI355X implementation of the kukan hot path.

The library must have been built (``python -m rspacedist_amd.build``); importing this
module on a machine without it raises instead of falling back to anything else.
Every entry point keeps the reference routine's keyword names (see ``abi.SIGNATURES``);
arrays may be numpy (host, Fortran order) or torch tensors (host or device).
"""
from __future__ import annotations

import ctypes
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkukan_hip.so")
_lib = None


class KukanError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KukanError(f"{LIB_PATH} is not built; run `python -m rspacedist_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.kukan_last_error.restype = ctypes.c_char_p
        L.kukan_init.argtypes = [ctypes.c_int]
        L.kukan_set_param.argtypes = [ctypes.c_char_p, ctypes.c_double]
        L.kukan_get_param.argtypes = [ctypes.c_char_p]
        L.kukan_get_param.restype = ctypes.c_double
        L.kukan_stream.restype = ctypes.c_void_p
        L.kukan_device_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.kukan_device_free.argtypes = [ctypes.c_void_p]
        L.kukan_memcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.kukan_fill_normal.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
        L.kukan_timing_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_double)]
        L.kukan_timing_tags.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.kukan_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.kukan_batch_add_system.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_int)]
        L.kukan_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.kukan_batch_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_int)]
        L.kukan_batch_destroy.argtypes = [ctypes.c_void_p]
        L.kukan_subspace_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_long,
                                          ctypes.c_void_p, ctypes.c_long, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_long]
        for f in ("kukan_bind",):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        for f in ("kukan_unbind", "kukan_sync_to_host", "kukan_sync_to_device"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        abi.bind(L, "kukan_")
        _lib = L
    return _lib


def last_error():
    return lib().kukan_last_error().decode()


def init(device=0):
    rc = lib().kukan_init(device)
    if rc:
        raise KukanError(last_error())


def set_param(key, value):
    lib().kukan_set_param(key.encode(), float(value))


def get_param(key):
    return lib().kukan_get_param(key.encode())


def call(name, **kw):
    """Call kukan_<name> with reference-named keyword arguments; raises KukanError with the
    library's message (the reference's stopp text) on a nonzero return code."""
    res = abi.call(lib(), "kukan_", name, **kw)
    if res["rc"] != 0:
        raise KukanError(f"kukan_{name} failed ({res['rc']}): {last_error()}")
    return res


def _ck(rc, what):
    if rc != 0:
        raise KukanError(f"{what} failed ({rc}): {last_error()}")


class DeviceArray:
    """A Fortran-ordered float64/int32 array in HBM, allocated through the library's HIP
    runtime.  Pass it to ``call`` like a numpy array; it is used in place."""

    def __init__(self, shape, dtype="float64"):
        import numpy as np
        self.shape = tuple(int(d) for d in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ck(lib().kukan_device_malloc(ctypes.byref(p), self.nbytes), "kukan_device_malloc")
        self.device_ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        _ck(lib().kukan_memcpy(d.device_ptr, a.ctypes.data, d.nbytes, 1), "kukan_memcpy")
        return d

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _ck(lib().kukan_memcpy(out.ctypes.data, self.device_ptr, self.nbytes, 2), "kukan_memcpy")
        _ck(lib().kukan_device_synchronize(), "kukan_device_synchronize")
        return out

    def fill_normal(self, seed):
        _ck(lib().kukan_fill_normal(self.device_ptr, self.nbytes // 8, seed), "kukan_fill_normal")
        return self

    def free(self):
        if self.device_ptr:
            _ck(lib().kukan_device_free(self.device_ptr), "kukan_device_free")
            self.device_ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def timing():
    """{tag: (ms, count, work)} accumulated since the last timing_reset()."""
    buf = ctypes.create_string_buffer(4096)
    _ck(lib().kukan_timing_tags(buf, 4096), "kukan_timing_tags")
    out = {}
    for tag in buf.value.decode().split(";"):
        if not tag:
            continue
        ms, cnt, work = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        _ck(lib().kukan_timing_get(tag.encode(), ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(work)),
            "kukan_timing_get")
        out[tag] = (ms.value, cnt.value, work.value)
    return out


def timing_reset():
    _ck(lib().kukan_timing_reset(), "kukan_timing_reset")


def synchronize():
    _ck(lib().kukan_device_synchronize(), "kukan_device_synchronize")


def subspace_gemm(trans_a, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """C(m x n) = alpha op(A) B + beta C on column-major arrays (DeviceArray or numpy)."""
    def p(x, name):
        return abi._ptr(x, "D", name)
    _ck(lib().kukan_subspace_gemm(int(trans_a), m, n, k, alpha, p(A, "A"), lda, p(B, "B"), ldb,
                                  beta, p(C, "C"), ldc), "kukan_subspace_gemm")


EXPORTS = ["kukan_init", "kukan_finalize", "kukan_last_error", "kukan_set_param",
           "kukan_get_param", "kukan_stream", "kukan_bind", "kukan_device_malloc", "kukan_device_free",
           "kukan_memcpy", "kukan_fill_normal", "kukan_device_synchronize", "kukan_timing_get",
           "kukan_timing_reset", "kukan_timing_tags", "kukan_unbind", "kukan_sync_to_host",
           "kukan_sync_to_device", "kukan_overlap_fdcheck_r", "kukan_rayleighritz_blacs_ini",
           "kukan_rayleighritz_blacs_fin", "kukan_subspace_gemm", "kukan_batch_create",
           "kukan_batch_add_system", "kukan_batch_step", "kukan_batch_get", "kukan_batch_destroy"] + ["kukan_" + n for n in abi.SIGNATURES]
