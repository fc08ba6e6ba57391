"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  See oracle.cpp for the
reference functions it restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.oracle_assemble.restype = C.c_int
        L.oracle_assemble.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)] + \
            [C.POINTER(vp)] * 6 + [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_free.argtypes = [vp]
        L.oracle_coefficients.restype = C.c_int
        L.oracle_coefficients.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_double)]
        L.oracle_lagrange2.restype = C.c_int
        L.oracle_lagrange2.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def lagrange2(elem_type, pts):
    """The oracle's order-2 geometry functions: array [4, nodes, npts] (None: no such type)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(4 * 27 * len(pts), np.float64)
    n = lib().oracle_lagrange2(elem_type, pts.ctypes.data_as(C.POINTER(C.c_double)), len(pts),
                               out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[:4 * n * len(pts)].reshape(4, n, len(pts)) if n else None


def coefficients(desc_ptr, numeric, bucket, n):
    """Coefficient table of one bucket: [(slot * nq + q) * n_elem + e]."""
    out = np.zeros(max(n, 1), np.float64)
    if lib().oracle_coefficients(desc_ptr, C.cast(C.byref(numeric), C.c_void_p), bucket,
                                 out.ctypes.data_as(C.POINTER(C.c_double))) != 0:
        raise RuntimeError("oracle_coefficients failed")
    return out[:n]


class OracleCSR:
    def __init__(self, Arow, Acol, Aval, Drow, Dcol, Dval, rhs, ke):
        self.Arow, self.Acol, self.Aval = Arow, Acol, Aval
        self.Drow, self.Dcol, self.Dval = Drow, Dcol, Dval
        self.rhs, self.ke = rhs, ke


def assemble(desc_ptr, numeric, ndofs, nfree, n_ke=None, nthreads=1):
    """Run the restated generate()+A()+b() on a plan description."""
    L = lib()
    nA, nD = C.c_int64(), C.c_int64()
    outs = [C.c_void_p() for _ in range(6)]
    rhs = np.zeros(max(ndofs, 1), np.float64)
    ke = np.zeros(max(n_ke, 1), np.float64) if n_ke else None
    rc = L.oracle_assemble(desc_ptr, C.cast(C.byref(numeric), C.c_void_p), nthreads, C.byref(nA), C.byref(nD),
                           *[C.byref(o) for o in outs], rhs.ctypes.data_as(C.POINTER(C.c_double)),
                           ke.ctypes.data_as(C.POINTER(C.c_double)) if ke is not None else None)
    if rc != 0:
        raise RuntimeError("oracle failed")

    def take(ptr, n, ctype, dtype):
        if n == 0:
            L.oracle_free(ptr)
            return np.zeros(0, dtype)
        a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()
        L.oracle_free(ptr)
        return a

    Arow = take(outs[0], nfree + 1, C.c_int64, np.int64)
    Acol = take(outs[1], nA.value, C.c_int32, np.int32)
    Aval = take(outs[2], nA.value, C.c_double, np.float64)
    Drow = take(outs[3], nfree + 1, C.c_int64, np.int64)
    Dcol = take(outs[4], nD.value, C.c_int32, np.int32)
    Dval = take(outs[5], nD.value, C.c_double, np.float64)
    return OracleCSR(Arow, Acol, Aval, Drow, Dcol, Dval, rhs[:ndofs], ke)
