"""ctypes loader of oracle/wmk_direct.c (the reference's direct wavemaker sums in C).

TEST INFRASTRUCTURE ONLY: used by the test suite as a checker and by bench.py's cpu_baseline
leg as the timed reference-mode wavemaker cost.  ``build()`` compiles it with gcc (OpenMP) into
oracle/libwmkdirect.so; ``load()`` builds it first when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "wmk_direct.c")
LIB = os.path.join(HERE, "libwmkdirect.so")
_lib = None


def build() -> None:
    subprocess.run(["gcc", "-O2", "-fopenmp", "-shared", "-fPIC", SRC, "-o", LIB, "-lm"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        lib = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        lib.phiadd_sl_rows.argtypes = [C.c_int] * 5 + [C.c_double] * 2 + [dp] * 8
        lib.phiadd_1st_rows.argtypes = [C.c_int] * 5 + [C.c_double] * 2 + [dp] * 3
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def phiadd_sl(n1, n2, xlen, ylen, a, da, eta, rows=None):
    """phiadd_SL's five fields (rows = (i1_lo, i1_hi), all rows by default)."""
    lib = load()
    lo, hi = rows if rows else (0, n1)
    a, da, eta = (np.asfortranarray(x, dtype=np.float64) for x in (a, da, eta))
    out = [np.zeros((n1, n2), order="F") for _ in range(5)]
    rc = lib.phiadd_sl_rows(n1, n2, a.shape[0], lo, hi, xlen, ylen, _p(a), _p(da), _p(eta), *[_p(o) for o in out])
    if rc:
        raise MemoryError("phiadd_sl_rows")
    return dict(zip(("phi", "phit", "phix", "phiy", "phiz"), out))


def phiadd_1st(n1, n2, xlen, ylen, a, da, rows=None):
    """phiadd_1st_SWEET's twelve z = 0 fields, shape (12, n1, n2)."""
    lib = load()
    lo, hi = rows if rows else (0, n1)
    a, da = (np.asfortranarray(x, dtype=np.float64) for x in (a, da))
    out = np.zeros((12, n2, n1))
    rc = lib.phiadd_1st_rows(n1, n2, a.shape[0], lo, hi, xlen, ylen, _p(a), _p(da), _p(out))
    if rc:
        raise MemoryError("phiadd_1st_rows")
    return np.transpose(out, (0, 2, 1))
