"""ctypes access to the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the CPU-baseline leg
of bench.py as the checker.  The product path (rspacedist_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

from rspacedist_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# routines only the oracle has: the one-system restatement of the batched band step
SIGNATURES = dict(abi.SIGNATURES)
SIGNATURES["band_sd_step"] = [("n", "i"), ("h", "d"), ("nbands", "i"), ("natom", "i"),
                              ("num_list", "i"), ("lstvec2", "I"), ("natinf", "I"),
                              ("vnlocp", "D"), ("dij", "D"), ("veff", "D"), ("svecre", "D"),
                              ("sval", "D")]
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.orc_last_error.restype = ctypes.c_char_p
        _lib.orc_dsyev.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        if hasattr(_lib, "orc_zheev"):
            _lib.orc_zheev.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        abi.bind(_lib, "orc_", SIGNATURES)
    return _lib


def last_error():
    return lib().orc_last_error().decode()


def call(name, **kw):
    """Call orc_<name> with reference-named keyword arguments; raises on a nonzero code."""
    return abi.call(lib(), "orc_", name, err=last_error, _table=SIGNATURES, **kw)


def dsyev(a):
    """Eigen-decomposition with the oracle's tred2/tql2 (ascending, canonical signs)."""
    import numpy as np
    z = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    w = np.zeros(z.shape[0])
    rc = lib().orc_dsyev(z.shape[0], z.ctypes.data, w.ctypes.data)
    if rc:
        raise RuntimeError(last_error())
    return w, z


def zheev(a):
    """Hermitian eigen-decomposition with the oracle's cyclic Jacobi (upper triangle read;
    ascending; each eigenvector's largest component real and positive)."""
    import numpy as np
    z = np.asfortranarray(a, dtype=np.complex128).copy(order="F")
    w = np.zeros(z.shape[0])
    rc = lib().orc_zheev(z.shape[0], z.ctypes.data, w.ctypes.data)
    if rc:
        raise RuntimeError(last_error())
    return w, z
