"""CPU checks of the include/fortfem_tet.h exports (registry and argument
checks; no device call): the reference C API's status codes
(include/fortfem.h, src/bindings/fortfem_capi_solver.f90:10-124)."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lib():
    from fortfem_amd import _capi

    return _capi.load()


def test_every_declared_function_is_exported():
    text = open(os.path.join(ROOT, "include", "fortfem_tet.h")).read()
    names = re.findall(r"^void\s+(fortfem_\w+)\s*\(", text, re.M)
    assert len(names) == 11
    L = lib()
    for n in names:
        assert hasattr(L, n), n


def _mesh(L):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
    t = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int32)
    h, st = C.c_int(0), C.c_int(9)
    L.fortfem_tetra_mesh_create(5, v.ctypes.data_as(C.c_void_p), 2, t.ctypes.data_as(C.c_void_p),
                                C.byref(h), C.byref(st))
    return h.value, st.value


def test_mesh_registry_has_32_slots_and_checks_handles():
    L = lib()
    handles = []
    for _ in range(32):
        h, st = _mesh(L)
        assert st == 0 and 1 <= h <= 32
        handles.append(h)
    assert len(set(handles)) == 32
    h, st = _mesh(L)
    assert st == -3 and h == 0  # registry exhausted
    st = C.c_int(9)
    for h in handles:
        L.fortfem_tetra_mesh_free(h, C.byref(st))
        assert st.value == 0
    L.fortfem_tetra_mesh_free(handles[0], C.byref(st))
    assert st.value == -3  # freed twice
    L.fortfem_tetra_mesh_free(0, C.byref(st))
    assert st.value == -3


def test_invalid_input_and_handles():
    L = lib()
    v = np.zeros((5, 3))
    t = np.array([[0, 1, 2, 5]], dtype=np.int32)  # vertex 5 out of range (zero-based)
    h, st = C.c_int(0), C.c_int(9)
    L.fortfem_tetra_mesh_create(5, v.ctypes.data_as(C.c_void_p), 1, t.ctypes.data_as(C.c_void_p),
                                C.byref(h), C.byref(st))
    assert st.value == -1 and h.value == 0
    n, nnz = C.c_int(0), C.c_int(0)
    L.fortfem_tetra_lagrange_stiffness_csc(7, 1, 2, C.c_double(1.0), C.c_double(0.0), 0, C.byref(n),
                                           C.byref(nnz), None, None, None, C.byref(st))
    assert st.value == -3  # no such mesh
    L.fortfem_tetra_lagrange_stiffness_csc(1, 9, 2, C.c_double(1.0), C.c_double(0.0), 0, C.byref(n),
                                           C.byref(nnz), None, None, None, C.byref(st))
    assert st.value == -1  # degree out of range
    it = C.c_int(0)
    x = np.zeros(3)
    L.fortfem_operator_solve(3, 3, C.c_double(1e-8), 10, 0, x.ctypes.data_as(C.c_void_p),
                             x.ctypes.data_as(C.c_void_p), C.byref(it), C.byref(st))
    assert st.value == -3
    L.fortfem_operator_free(3, C.byref(st))
    assert st.value == -3
    cp = np.array([0, 1, 3], dtype=np.int32)  # col_ptr[n] != nnz
    ri = np.array([0, 1], dtype=np.int32)
    va = np.ones(2)
    L.fortfem_real_operator_csc(2, 2, cp.ctypes.data_as(C.c_void_p), ri.ctypes.data_as(C.c_void_p),
                                va.ctypes.data_as(C.c_void_p), C.byref(h), C.byref(st))
    assert st.value == -1 and h.value == 0


def test_complex_operator_argument_checks():
    """fortfem_complex_operator_csc validates the zero-based CSC like
    valid_csc_input (fortfem_capi_solver.f90:126-142) before any device
    call; a complex handle is refused by the real solve and vice versa."""
    L = lib()
    cp = np.array([0, 1, 2], dtype=np.int32)
    ri = np.array([0, 2], dtype=np.int32)  # row out of range
    va = np.ones(4)
    h, st = C.c_int(7), C.c_int(9)
    L.fortfem_complex_operator_csc(2, 2, cp.ctypes.data_as(C.c_void_p), ri.ctypes.data_as(C.c_void_p),
                                   va.ctypes.data_as(C.c_void_p), C.byref(h), C.byref(st))
    assert st.value == -1 and h.value == 0
    it = C.c_int(0)
    x = np.zeros(4)
    L.fortfem_complex_operator_solve(5, 2, C.c_double(1e-10), 10, va.ctypes.data_as(C.c_void_p),
                                     x.ctypes.data_as(C.c_void_p), C.byref(it), C.byref(st))
    assert st.value == -3  # unknown handle


def test_complex_operator_refuses_nonsymmetric_matrix():
    """COCG needs A = A^T: a complex CSC whose (0, 1) and (1, 0) entries
    differ is refused with -1 at upload, before any device call; the same
    pattern with equal off-diagonal entries passes the check (and then
    reaches the device, which is absent here: -4, or 0 on a GPU host)."""
    L = lib()
    cp = np.array([0, 2, 4], dtype=np.int32)
    ri = np.array([0, 1, 0, 1], dtype=np.int32)
    # column-major (re, im): A00 = 2, A10 = 1+1j, A01 = 1-1j, A11 = 3
    va = np.array([2, 0, 1, 1, 1, -1, 3, 0], dtype=np.float64)
    h, st = C.c_int(7), C.c_int(9)
    L.fortfem_complex_operator_csc(2, 4, cp.ctypes.data_as(C.c_void_p), ri.ctypes.data_as(C.c_void_p),
                                   va.ctypes.data_as(C.c_void_p), C.byref(h), C.byref(st))
    assert st.value == -1 and h.value == 0
    va[4:6] = [1, 1]  # A01 = A10: complex symmetric (not Hermitian)
    L.fortfem_complex_operator_csc(2, 4, cp.ctypes.data_as(C.c_void_p), ri.ctypes.data_as(C.c_void_p),
                                   va.ctypes.data_as(C.c_void_p), C.byref(h), C.byref(st))
    assert st.value in (0, -4)
    if st.value == 0:
        L.fortfem_operator_free(h.value, C.byref(st))
