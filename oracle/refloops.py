"""ctypes wrapper of oracle/librefloops.so, the C restatement of the
reference CPU path with its own loop nests (oracle/refloops.c).  Test and
measurement infrastructure only: tests pin it against ``oracle.ccsd`` and
bench.py times it as the CPU baseline.  Build: ``make -C oracle``.

Arrays at this interface are numpy arrays in the oracle's convention (C, F,
G n x n; t1[i,a], t2[i,j,a,b] with virtual indices from 0); they are passed
to the library in the reference's Fortran order.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefloops.so")
_P = ctypes.POINTER(ctypes.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built; run make -C oracle")
        lib = ctypes.CDLL(LIB_PATH)
        i, d = ctypes.c_int, ctypes.c_double
        lib.ref_build_R.argtypes = [i, _P, _P, _P]
        for name in ("ref_singles", "ref_doubles"):
            getattr(lib, name).argtypes = [i, i, _P, _P, _P, _P, _P]
        lib.ref_solve.restype = i
        lib.ref_solve.argtypes = [i, i, _P, _P, _P, d, d, i, i, i, _P, ctypes.POINTER(i), _P, _P]
        lib.ref_time_parts.restype = i
        lib.ref_time_parts.argtypes = [i, i, _P, _P, _P, i, i, i, _P, _P]
        _lib = lib
    return _lib


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return a.ctypes.data_as(_P)


def build_R(C, G):
    n = C.shape[0]
    Cf, Gf = _f(C), _f(G)
    R = np.zeros((n, n, n, n), order="F")
    load().ref_build_R(n, _p(Cf), _p(Gf), _p(R))
    return np.ascontiguousarray(R)


def residuals(C, G, F, o, t1, t2):
    """(d1[i,a], d2[i,j,a,b]) of proj-ccsd-spat-hf-1/2.f90 at the given amplitudes."""
    n = C.shape[0]
    v = n - o
    lib = load()
    Cf, Gf, Ff = _f(C), _f(G), _f(F)
    R = np.zeros((n, n, n, n), order="F")
    lib.ref_build_R(n, _p(Cf), _p(Gf), _p(R))
    a1, a2 = _f(t1), _f(t2)
    d1 = np.zeros((o, v), order="F")
    d2 = np.zeros((o, o, v, v), order="F")
    lib.ref_singles(n, o, _p(Ff), _p(R), _p(a1), _p(a2), _p(d1))
    lib.ref_doubles(n, o, _p(Ff), _p(R), _p(a1), _p(a2), _p(d2))
    return np.ascontiguousarray(d1), np.ascontiguousarray(d2)


def solve(C, F, G, o, eref, eps=1e-10, maxiter=2000, diis_steps=20, threads=1):
    n = C.shape[0]
    v = n - o
    Cf, Gf, Ff = _f(C), _f(G), _f(F)
    e = np.zeros(5)
    it = ctypes.c_int()
    t1 = np.zeros((o, v), order="F")
    t2 = np.zeros((o, o, v, v), order="F")
    st = load().ref_solve(n, o, _p(Cf), _p(Ff), _p(Gf), float(eref), float(eps), int(maxiter), int(diis_steps),
                          int(threads), _p(e), ctypes.byref(it), _p(t1), _p(t2))
    return dict(status=st, iterations=it.value, energy=e, t1=np.ascontiguousarray(t1), t2=np.ascontiguousarray(t2))


def time_parts(C, F, G, o, iters=1, diis_steps=20, threads=1):
    """(seconds of the R build, seconds of `iters` iterations from the guess)."""
    n = C.shape[0]
    Cf, Gf, Ff = _f(C), _f(G), _f(F)
    tb = ctypes.c_double()
    ti = ctypes.c_double()
    rc = load().ref_time_parts(n, o, _p(Cf), _p(Ff), _p(Gf), int(iters), int(diis_steps), int(threads),
                               ctypes.byref(tb), ctypes.byref(ti))
    if rc != 0:
        raise MemoryError("ref_time_parts: allocation failed")
    return tb.value, ti.value
