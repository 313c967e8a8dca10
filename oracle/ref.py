"""ctypes access to oracle/_ref/libffo_ref.so: reference modules compiled from
their own sources (oracle/ref.mk).  TEST INFRASTRUCTURE ONLY -- used by the
CPU tests to pin the oracle; absent (and skipped) where the reference sources
are not mounted."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libffo_ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lagrange_eval(degree, point):
    n = (degree + 1) * (degree + 2) * (degree + 3) // 6
    v = np.zeros(n)
    g = np.zeros((n, 3))
    pt = np.ascontiguousarray(point, dtype=np.float64)
    st = lib().ffr_lagrange_eval(C.c_int(degree), _p(pt), _p(v), _p(g))
    return st, v, g


def lagrange_dof_map(degree, tets):
    t = np.ascontiguousarray(tets, dtype=np.int32)
    n = max((degree + 1) * (degree + 2) * (degree + 3) // 6, 1)
    g = np.zeros((t.shape[0], n), dtype=np.int32)
    cnt = C.c_int(0)
    st = lib().ffr_lagrange_dof_map(C.c_int(degree), C.c_int(t.shape[0]), _p(t), _p(g), C.byref(cnt))
    return st, g, cnt.value


def edge_dof_map(tets):
    t = np.ascontiguousarray(tets, dtype=np.int32)
    T = t.shape[0]
    edges = np.zeros((6 * T, 2), dtype=np.int32)
    gd = np.zeros((T, 6), dtype=np.int32)
    orient = np.zeros((T, 6), dtype=np.int32)
    ne = C.c_int(0)
    st = lib().ffr_edge_dof_map(C.c_int(T), _p(t), C.byref(ne), _p(edges), _p(gd), _p(orient))
    return st, edges[:ne.value].copy(), gd, orient


def whitney(point):
    v = np.zeros((6, 3))
    c = np.zeros((6, 3))
    pt = np.ascontiguousarray(point, dtype=np.float64)
    st = lib().ffr_whitney(_p(pt), _p(v), _p(c))
    return st, v, c


def nedelec_basis(order, point):
    n = order * (order + 2) * (order + 3) // 2
    v = np.zeros((n, 3))
    c = np.zeros((n, 3))
    pt = np.ascontiguousarray(point, dtype=np.float64)
    st = lib().ffr_nedelec_basis(C.c_int(order), _p(pt), _p(v), _p(c))
    return st, v, c


def face_basis_to_local(order, perm):
    """the k(k-1) square matrix basis_to_local(perm) column by column"""
    n = order * (order - 1)
    p = np.ascontiguousarray(perm, dtype=np.int32)
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        out = np.zeros(n)
        st = lib().ffr_face_basis_to_local(C.c_int(order), _p(p), _p(e), _p(out))
        if st != 0:
            raise ValueError("face transform failed")
        M[:, j] = out
    return M


def h1_oracle(x, y, z):
    v = np.zeros(5)
    lib().ffr_h1_oracle(C.c_double(x), C.c_double(y), C.c_double(z), _p(v))
    return v
