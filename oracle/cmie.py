"""ctypes loader of liboracle_mie.so (built from cmie.c by oracle/Makefile).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_mie.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_mie.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.cmie_batch.restype = ctypes.c_int
        L.cmie_batch.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def batch(x, m, y=None, m2=None, threads=0, stable=False):
    """props [6, n] (qext, qsca, qabs, qb, asy, qratio) for n points, mu = 1.

    stable=False: the reference's literal coated formula; stable=True: the
    cancellation-free rearrangement of the same coefficients (cmie.c).
    """
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    mm = np.ascontiguousarray(np.broadcast_to(np.asarray(m, np.complex128), (n,))).view(np.float64)
    out = np.empty(6 * n)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    m2v = None if m2 is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(m2, np.complex128), (n,))).view(np.float64)
    p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    lib().cmie_batch(n, p(x), p(yy), p(mm), p(m2v), p(out), int(threads), int(bool(stable)))
    return out.reshape(6, n)
