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


def chain_homog(x, m, r, row, w, n_rows, cost, threads=0):
    """Homogeneous runoptics chain on the host (cmie.c cmie_chain_homog).

    x [n] size parameters, m index (scalar or [n]), r [n] radii, row [n] the
    weight row of each point, w [n] its number weight, cost [n_ang] cosines.
    Returns (pm [n_rows, 4, n_ang] weighted sums of s11, s12, s33, s34;
    sc [n_rows, 4] = sum w pi r^2, ... * Qext, ... * Qsca, ... * Qsca g).
    """
    from . import mie as M
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    mm = np.ascontiguousarray(np.broadcast_to(np.asarray(m, np.complex128), (n,))).view(np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    row = np.ascontiguousarray(row, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    cost = np.asarray(cost, dtype=np.float64)
    cap = int(M.nmax_of(float(x.max())))
    P, T = M.pt_table(cost, cap)                       # [n_ang, cap]
    pt = np.ascontiguousarray(np.concatenate([P.T.ravel(), T.T.ravel()]))
    pm = np.empty(n_rows * 4 * cost.size)
    sc = np.empty(n_rows * 4)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    f = lib().cmie_chain_homog
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int,
                                                            ctypes.c_void_p, ctypes.c_int,
                                                            ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.c_int]
    rc = f(n, p(x), p(mm), p(r), p(row), p(w), int(n_rows), int(cost.size), p(pt), cap, p(pm),
           p(sc), int(threads))
    if rc != 0:
        raise ValueError("cmie_chain_homog: a point exceeds the table length")
    return pm.reshape(n_rows, 4, cost.size), sc.reshape(n_rows, 4)
