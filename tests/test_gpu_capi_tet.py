"""include/fortfem_tet.h on the GPU, called through ctypes as a C host would.

CSC exports (zero-based, the capacity protocol of fortfem_capi_mesh.f90:
545-571) against the oracle's assemblies: pattern bit-exact, values 1e-12;
retained operators (mesh-assembled with boundary elimination, or uploaded
CSC) solved by Jacobi-PCG against the oracle's reduction and PCG to 1e-8;
the 32-slot operator registry (fortfem_capi_solver.f90:10-70) and its
status codes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from fortfem_amd import _capi

    return _capi.load()


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def mesh_handle(L, v, t0):
    h, st = C.c_int(0), C.c_int(9)
    L.fortfem_tetra_mesh_create(v.shape[0], p(np.ascontiguousarray(v)), t0.shape[0],
                                p(np.ascontiguousarray(t0, dtype=np.int32)), C.byref(h), C.byref(st))
    assert st.value == 0
    return h.value


def export(L, fn, args):
    n, nnz, st = C.c_int(0), C.c_int(0), C.c_int(9)
    fn(*args, 0, C.byref(n), C.byref(nnz), None, None, None, C.byref(st))
    assert st.value == -2 and n.value > 0 and nnz.value > 0  # capacity too small, sizes returned
    cp = np.zeros(n.value + 1, dtype=np.int32)
    ri = np.zeros(nnz.value, dtype=np.int32)
    va = np.zeros(nnz.value)
    fn(*args, nnz.value, C.byref(n), C.byref(nnz), p(cp), p(ri), p(va), C.byref(st))
    assert st.value == 0
    return cp, ri, va


def same(cp, ri, va, a):
    assert np.array_equal(cp, a.col_ptr - 1) and np.array_equal(ri, a.row_idx - 1)
    assert np.max(np.abs(va - a.val)) <= 1e-12 * np.max(np.abs(a.val))


def test_lagrange_and_nedelec_exports(L):
    v, t = O.unit_box_mesh(3)
    h = mesh_handle(L, v, t - 1)
    cp, ri, va = export(L, L.fortfem_tetra_lagrange_stiffness_csc, (h, 2, 4, C.c_double(1.3), C.c_double(0.2)))
    a, _, _ = O.lagrange_stiffness_csc(v, t, 2, 4, kappa=1.3, mu=0.2)
    same(cp, ri, va, a)
    cp, ri, va = export(L, L.fortfem_tetra_nedelec_curl_mass_csc, (h, 2, C.c_double(0.7), C.c_double(1.1)))
    same(cp, ri, va, O.nedelec_csc(2, v, t, alpha=0.7, beta=1.1))
    st = C.c_int(9)
    L.fortfem_tetra_mesh_free(h, C.byref(st))
    assert st.value == 0


def test_mesh_operator_with_boundary_matches_oracle(L):
    v, t = O.unit_box_mesh(4)
    h = mesh_handle(L, v, t - 1)
    op, n, st = C.c_int(0), C.c_int(0), C.c_int(9)
    L.fortfem_tetra_lagrange_operator(h, 2, 6, C.c_double(1.0), C.c_double(0.0), 1, C.byref(op), C.byref(n),
                                      C.byref(st))
    assert st.value == 0 and op.value > 0
    a, g, G = O.lagrange_stiffness_csc(v, t, 2, 6)
    assert n.value == G
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    b = np.random.default_rng(1).uniform(-1, 1, G)
    red, free, rb = O.constrained_reduction(a, cons, np.zeros(G), b)
    xo, sto = O.solve_sparse(red, rb, method="pcg", preconditioner="jacobi", tolerance=1e-12,
                             max_iterations=10000, breakdown="relative")
    for prec in (0, 1):
        x = np.zeros(G)
        it = C.c_int(0)
        L.fortfem_operator_solve(op.value, G, C.c_double(1e-12), 10000, prec, p(b), p(x), C.byref(it),
                                 C.byref(st))
        assert st.value == 0 and it.value > 0
        assert np.linalg.norm(x[free - 1] - xo) <= 1e-8 * np.linalg.norm(xo)
        assert not x[np.asarray(cons, dtype=bool)].any()
    x = np.zeros(G)
    L.fortfem_operator_solve(op.value, G, C.c_double(1e-12), 3, 0, p(b), p(x), C.byref(it), C.byref(st))
    assert st.value == 2 and it.value == 3  # iteration limit
    L.fortfem_operator_solve(op.value, G + 1, C.c_double(1e-12), 10, 0, p(b), p(x), C.byref(it), C.byref(st))
    assert st.value == -1  # size mismatch
    L.fortfem_operator_free(op.value, C.byref(st))
    assert st.value == 0
    L.fortfem_tetra_mesh_free(h, C.byref(st))


def test_nedelec_operator_zero_tangential(L):
    v, t = O.unit_box_mesh(2)
    h = mesh_handle(L, v, t - 1)
    op, n, st = C.c_int(0), C.c_int(0), C.c_int(9)
    L.fortfem_tetra_nedelec_operator(h, 2, C.c_double(1.0), C.c_double(1.0), 1, C.byref(op), C.byref(n),
                                     C.byref(st))
    assert st.value == 0
    a = O.nedelec_csc(2, v, t)
    G = a.n_rows
    assert n.value == G
    b = np.random.default_rng(2).uniform(-1, 1, G)
    x = np.zeros(G)
    it = C.c_int(0)
    L.fortfem_operator_solve(op.value, G, C.c_double(1e-12), 20000, 0, p(b), p(x), C.byref(it), C.byref(st))
    assert st.value == 0
    # the constrained (boundary) entries are zero and the free ones solve the
    # free block of the oracle's matrix
    mask = x == 0.0
    A = a.to_dense()
    free = ~mask
    r = A[np.ix_(free, free)] @ x[free] - b[free]
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b[free])
    L.fortfem_operator_free(op.value, C.byref(st))
    L.fortfem_tetra_mesh_free(h, C.byref(st))


def test_csc_operator_registry(L):
    v, t = O.unit_box_mesh(2)
    a, _, G = O.lagrange_stiffness_csc(v, t, 1, 2, mu=1.0)
    cp, ri, va = a.col_ptr - 1, a.row_idx - 1, a.val
    handles = []
    st = C.c_int(9)
    for _ in range(32):
        h = C.c_int(0)
        L.fortfem_real_operator_csc(G, a.nnz, p(cp), p(ri), p(va), C.byref(h), C.byref(st))
        assert st.value == 0
        handles.append(h.value)
    assert sorted(handles) == list(range(1, 33))
    h = C.c_int(0)
    L.fortfem_real_operator_csc(G, a.nnz, p(cp), p(ri), p(va), C.byref(h), C.byref(st))
    assert st.value == -3 and h.value == 0  # registry full
    b = np.linspace(-1, 1, G)
    xo, _ = O.solve_sparse(a, b, method="pcg", preconditioner="jacobi", tolerance=1e-12, max_iterations=1000,
                           breakdown="relative")
    x = np.zeros(G)
    it = C.c_int(0)
    L.fortfem_operator_solve(handles[5], G, C.c_double(1e-12), 1000, 0, p(b), p(x), C.byref(it), C.byref(st))
    assert st.value == 0 and np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    for hh in handles:
        L.fortfem_operator_free(hh, C.byref(st))
        assert st.value == 0
    L.fortfem_operator_free(handles[0], C.byref(st))
    assert st.value == -3


def _complex_operator(L, a):
    cp, ri = a.col_ptr - 1, a.row_idx - 1
    va = np.ascontiguousarray(np.stack([a.val.real, a.val.imag], axis=-1).reshape(-1))
    h, st = C.c_int(0), C.c_int(9)
    L.fortfem_complex_operator_csc(a.n_rows, a.nnz, p(cp.astype(np.int32)), p(ri.astype(np.int32)), p(va),
                                   C.byref(h), C.byref(st))
    assert st.value == 0 and h.value > 0
    return h.value


def _complex_solve(L, h, n, b, tol=1e-12, maxit=20000):
    bb = np.ascontiguousarray(np.stack([b.real, b.imag], axis=-1).reshape(-1))
    x = np.zeros(2 * n)
    it, st = C.c_int(0), C.c_int(9)
    L.fortfem_complex_operator_solve(h, n, C.c_double(tol), maxit, p(bb), p(x), C.byref(it), C.byref(st))
    return x[0::2] + 1j * x[1::2], it.value, st.value


def test_complex_operator_cocg_matches_dense_solve(L):
    """Complex-symmetric operators (the PML family, and curl-curl with a
    complex mass shift) retained in place of fortfem_complex_factor_csc and
    solved by Jacobi-COCG (fortfem_complex_solve solves directly): the
    solution within 1e-8 of a dense complex solve of the same matrix."""
    v, t = O.unit_box_mesh(2)
    rng = np.random.default_rng(5)
    cases = []
    a = O.nedelec_csc(2, v, t, alpha=1.0, beta=1.0)
    m = O.nedelec_csc(2, v, t, alpha=0.0, beta=1.0)
    assert np.array_equal(a.row_idx, m.row_idx)
    a.val = a.val + 1j * m.val
    cases.append(a)
    s = rng.uniform(0.8, 1.2, (t.shape[0], 3)) + 1j * rng.uniform(0.2, 0.6, (t.shape[0], 3))
    cases.append(O.nedelec_pml_csc(1, v, t, s, 0.7))
    for a in cases:
        A = np.zeros((a.n_rows, a.n_cols), dtype=complex)
        for c in range(a.n_cols):
            for k in range(a.col_ptr[c] - 1, a.col_ptr[c + 1] - 1):
                A[a.row_idx[k] - 1, c] += a.val[k]
        assert np.max(np.abs(A - A.T)) <= 1e-13 * np.max(np.abs(A))
        b = rng.uniform(-1, 1, a.n_rows) + 1j * rng.uniform(-1, 1, a.n_rows)
        xr = np.linalg.solve(A, b)
        h = _complex_operator(L, a)
        x, it, st = _complex_solve(L, h, a.n_rows, b)
        assert st == 0 and it > 0
        assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
        assert np.linalg.norm(A @ x - b) <= 1e-11 * np.linalg.norm(b)
        # iteration limit, size mismatch, a complex handle in the real solve
        x, it, st = _complex_solve(L, h, a.n_rows, b, maxit=2)
        assert st == 2 and it == 2
        _, _, st = _complex_solve(L, h, a.n_rows + 1, b)
        assert st == -1
        rb = np.zeros(a.n_rows)
        xx = np.zeros(a.n_rows)
        it, stc = C.c_int(0), C.c_int(9)
        L.fortfem_operator_solve(h, a.n_rows, C.c_double(1e-10), 10, 0, p(rb), p(xx), C.byref(it), C.byref(stc))
        assert stc.value == -1
        L.fortfem_operator_free(h, C.byref(stc))
        assert stc.value == 0
