"""ctypes wrapper of ``oracle/brute.c`` (test infrastructure only)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libpvdoracle.so")
_lib = None


def build() -> str:
    """Compile ``brute.c`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "libpvdoracle.so"], check=True)
    return _LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        lib = ctypes.CDLL(_LIB)
        f = lib.pvd_oracle_brute_conv3d
        f.restype = None
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_void_p]
        _lib = lib
    return _lib


def brute_conv3d(activity: np.ndarray, kernel: np.ndarray, mode: str = "reference") -> np.ndarray:
    """Direct convolution in C: ``mode='reference'`` (circular, origin-anchored,
    cropped kernel) or ``'linear'`` (centred 'same', zero outside)."""
    a = np.ascontiguousarray(activity, dtype=np.float64)
    k = np.ascontiguousarray(kernel, dtype=np.float64)
    out = np.empty_like(a)
    _load().pvd_oracle_brute_conv3d(a.ctypes.data, *a.shape, k.ctypes.data, *k.shape,
                                    0 if mode == "reference" else 1, out.ctypes.data)
    return out
