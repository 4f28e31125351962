"""ctypes front end of the C banded restatement (``oracle/pade_banded.c``).

TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's CPU-baseline leg).

The band rows of A and the stencil / boundary rows of B are extracted from the
dense assembly in ``pade_dense`` so the C code checks the factorisation and the
sweep order of the reference algorithm, while the coefficients come from one
place.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import pade_dense

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_pade.so")
_lib = None

MAXK = 9


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        lib = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp,
                                      ctypes.c_int, dp, ctypes.c_int, dp]
        lib.oracle_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_apply.restype = ctypes.c_int
        lib.oracle_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, ctypes.c_long,
                                     ctypes.c_long, ctypes.c_int]
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tables(kind, which, n, dx, periodic, bc1, bcn):
    """Band rows of A, interior stencil and boundary rows of B (1/dx folded)."""
    A, B = pade_dense.assemble(kind, which, n, dx, periodic, bc1, bcn)
    w = 1 if kind == "cd06" else 2
    band = np.zeros((n, 2 * w + 1))
    for i in range(n):
        for o in range(-w, w + 1):
            j = (i + o) % n if periodic else i + o
            if 0 <= j < n:
                band[i, o + w] = A[i, j]
    Aint, Bint = pade_dense.interior(kind, which)
    H = max(abs(o) for o in Bint)
    st = np.zeros(9)
    s = dx ** (-pade_dense.scale_power(kind, which))
    for o, v in Bint.items():
        st[o + 4] = v * s
    nb = 0
    if not periodic:
        nb = {"cd06": 3, "cd10": 4 if which == "d1" else 3, "cf90": 4}[kind]
    bl = np.zeros((4, MAXK))
    br = np.zeros((4, MAXK))
    for q in range(nb):
        for c in range(min(MAXK, n)):
            bl[q, c] = B[q, c]
            br[q, c] = B[n - 1 - q, n - 1 - c]
    return band, H, st, nb, bl, br


class BandedOp:
    """One operator instance of the C restatement."""

    def __init__(self, kind, which, n, dx=1.0, periodic=True, bc1=0, bcn=0):
        lib = _load()
        self.n = n
        band, H, st, nb, bl, br = tables(kind, which, n, dx, periodic, bc1, bcn)
        self._keep = (np.ascontiguousarray(band), st, bl, br)
        w = band.shape[1] // 2
        self.h = lib.oracle_create(n, w, int(periodic), _dp(self._keep[0]), H, _dp(st), nb,
                                   _dp(bl), nb, _dp(br))
        if not self.h:
            raise RuntimeError("oracle factorisation failed")

    def apply(self, f, axis, nthreads=1, out=None):
        """f: Fortran-ordered float64 array (n1, n2, n3); axis 0/1/2."""
        lib = _load()
        if not (f.flags.f_contiguous and f.dtype == np.float64):
            f = np.asfortranarray(f, dtype=np.float64)
        if out is None:
            out = np.empty_like(f, order="F")
        n1, n2, n3 = f.shape
        if axis == 0:
            na, nb = n2, n3
        elif axis == 1:
            na, nb = n1, n3
        else:
            na, nb = n1, n2
        assert f.shape[axis] == self.n
        rc = lib.oracle_apply(self.h, axis + 1, _dp(f), _dp(out), na, nb, nthreads)
        if rc != 0:
            raise RuntimeError("oracle_apply failed")
        return out

    def __del__(self):
        try:
            if self.h:
                _load().oracle_destroy(self.h)
                self.h = None
        except Exception:
            pass


def max_threads():
    return _load().oracle_max_threads()
