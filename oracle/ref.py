"""ctypes access to the reference's own Fortran routines (oracle/_ref/libkukan_ref.so, built
by `make -C oracle ref` from the reference sources where they lie).

TEST INFRASTRUCTURE ONLY: the tests and the golden-vector script use it to check the oracle
restatement against the reference itself; nothing else loads it.  The library exists only
where the reference sources were present at build time (not on the GPU boxes).
"""
from __future__ import annotations

import ctypes
import os

from rspacedist_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = abi.SIGNATURES
LIB = os.path.join(HERE, "_ref", "libkukan_ref.so")
_lib = None


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{LIB} is not built (`make -C oracle ref` needs the reference sources)")
        # global symbol scope: MKL's core library opens its CPU-specific kernels at run time,
        # and they resolve symbols of the MKL threading layer this library links
        L = ctypes.CDLL(LIB, mode=ctypes.RTLD_GLOBAL)
        L.ref_tools_definefdcoef.argtypes = [ctypes.c_int, ctypes.c_void_p]
        abi.bind(L, "ref_", TABLE)
        if L.ref_init() != 0:
            raise RuntimeError("ref_init failed")
        _lib = L
    return _lib


def call(name, **kw):
    """Call the reference routine `name` with reference-named keyword arguments."""
    return abi.call(lib(), "ref_", name, _table=TABLE, **kw)


def definefdcoef(nf):
    import numpy as np
    a = np.zeros(9)
    lib().ref_tools_definefdcoef(nf, a.ctypes.data)
    return a
