"""FP64 CPU oracle for the fortfem sparse-assembly + Jacobi-PCG path.

TEST INFRASTRUCTURE ONLY.  This package wraps ``oracle/libffo.so`` (a plain-C
restatement of the reference algorithms; every C function cites the reference
file:line it follows).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline.  The product package ``fortfem_amd`` never imports it.

Array conventions mirror the reference's Fortran storage: ``vertices`` is
(nv, 3) C-order == vertices(3, nv) column-major, ``tetrahedra`` is (T, 4) with
1-based ids, ``global_dofs`` is (T, n_loc) with 1-based ids, CSC arrays are
1-based like fortsparse's ``csc_t``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = C.POINTER(C.c_int64)


def lib_path() -> str:
    return os.path.join(_HERE, "libffo.so")


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "libffo.so"], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        L = _LIB
        L.ffo_det3.restype = C.c_double
        L.ffo_det3.argtypes = [_dp]
    return _LIB


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _check(status, what):
    if status != 0:
        raise ValueError(f"oracle {what} failed with status {status}")


# ---------------------------------------------------------------- meshes ---
def box_mesh(bounds, counts):
    """generate_structured_tetra_box_mesh; bounds = [[lo..], [hi..]]."""
    b = _f(np.asarray(bounds, dtype=np.float64).reshape(2, 3).reshape(-1))
    c = _i(counts)
    nv = int(np.prod(c + 1))
    nt = 6 * int(np.prod(c))
    v = np.zeros((nv, 3))
    t = np.zeros((nt, 4), dtype=np.int32)
    _check(_lib().ffo_box_mesh(b.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                               v.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p)),
           "box_mesh")
    return v, t


def unit_box_mesh(n):
    return box_mesh([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], [n, n, n])


def unit_square_mesh(n):
    v = np.zeros((n * n, 2))
    t = np.zeros((2 * (n - 1) ** 2, 3), dtype=np.int32)
    bnd = np.zeros(n * n, dtype=np.int32)
    _check(_lib().ffo_unit_square_mesh(C.c_int(n), v.ctypes.data_as(C.c_void_p),
                                       t.ctypes.data_as(C.c_void_p),
                                       bnd.ctypes.data_as(C.c_void_p)), "unit_square_mesh")
    return v, t, bnd.astype(bool)


# ------------------------------------------------------------ quadrature ---
def gauss_legendre_ab(n, a, b):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(_lib().ffo_gauss_legendre_ab(C.c_int(n), C.c_double(a), C.c_double(b),
                                        x.ctypes.data_as(C.c_void_p),
                                        w.ctypes.data_as(C.c_void_p)), "gauss_legendre")
    return x, w


def tetra_duffy(p):
    nq = _lib().ffo_tetra_duffy_count(C.c_int(p))
    if nq <= 0:
        raise ValueError("negative degree")
    x, y, z, w = (np.zeros(nq) for _ in range(4))
    _check(_lib().ffo_tetra_duffy(C.c_int(p), *(a.ctypes.data_as(C.c_void_p) for a in (x, y, z, w))),
           "tetra_duffy")
    return x, y, z, w


def triangle_duffy(p):
    nq = _lib().ffo_triangle_duffy_count(C.c_int(p))
    if nq <= 0:
        raise ValueError("negative degree")
    x, y, w = (np.zeros(nq) for _ in range(3))
    _check(_lib().ffo_triangle_duffy(C.c_int(p), *(a.ctypes.data_as(C.c_void_p) for a in (x, y, w))),
           "triangle_duffy")
    return x, y, w


# ------------------------------------------------------------- Lagrange ---
def lagrange_dof_count(d):
    return _lib().ffo_lagrange_dof_count(C.c_int(d))


def lagrange_init(d):
    n = lagrange_dof_count(d)
    bary = np.zeros((n, 4), dtype=np.int32)
    nodes = np.zeros((n, 3))
    _check(_lib().ffo_lagrange_init(C.c_int(d), bary.ctypes.data_as(C.c_void_p),
                                    nodes.ctypes.data_as(C.c_void_p)), "lagrange_init")
    return bary, nodes


def lagrange_eval(d, point):
    n = lagrange_dof_count(d)
    p = _f(point)
    vals = np.zeros(n)
    grads = np.zeros((n, 3))
    st = _lib().ffo_lagrange_eval(C.c_int(d), p.ctypes.data_as(C.c_void_p),
                                  vals.ctypes.data_as(C.c_void_p), grads.ctypes.data_as(C.c_void_p))
    return st, vals, grads


def lagrange_dof_map(d, tets, search=False):
    t = _i(tets)
    n = max(lagrange_dof_count(d), 1)
    g = np.zeros((t.shape[0], n), dtype=np.int32)
    cnt = C.c_int(0)
    fn = _lib().ffo_lagrange_dof_map_search if search else _lib().ffo_lagrange_dof_map
    st = fn(C.c_int(d), C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p),
            g.ctypes.data_as(C.c_void_p), C.byref(cnt))
    return st, g, cnt.value


def lagrange_stiffness_element(vertices4, d, q, kappa=1.0, mu=0.0):
    v = _f(vertices4)
    n = lagrange_dof_count(d)
    m = np.zeros((n, n))  # m[c, r] in C order == matrix(r, c) column-major
    st = _lib().ffo_lagrange_stiffness_element(v.ctypes.data_as(C.c_void_p), C.c_int(d), C.c_int(q),
                                               C.c_double(kappa), C.c_double(mu),
                                               m.ctypes.data_as(C.c_void_p))
    return st, m.T.copy()


@dataclass
class CSC:
    n_rows: int
    n_cols: int
    col_ptr: np.ndarray  # 1-based int32
    row_idx: np.ndarray  # 1-based int32
    val: np.ndarray

    @property
    def nnz(self):
        return int(self.row_idx.shape[0])

    def to_dense(self):
        a = np.zeros((self.n_rows, self.n_cols))
        for c in range(self.n_cols):
            lo, hi = self.col_ptr[c] - 1, self.col_ptr[c + 1] - 1
            a[self.row_idx[lo:hi] - 1, c] += self.val[lo:hi]
        return a

    def matvec(self, x):
        return csc_matvec(self, x)


def csc_from_triplet(nrow, ncol, rows, cols, vals):
    r, c, v = _i(rows), _i(cols), _f(vals)
    nt = r.shape[0]
    cp = np.zeros(ncol + 1, dtype=np.int32)
    ri = np.zeros(max(nt, 1), dtype=np.int32)
    va = np.zeros(max(nt, 1))
    nnz = C.c_int64(0)
    _check(_lib().ffo_csc_from_triplet(C.c_int(nrow), C.c_int(ncol), C.c_int64(nt),
                                       r.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                                       v.ctypes.data_as(C.c_void_p), cp.ctypes.data_as(C.c_void_p),
                                       ri.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                       C.byref(nnz)), "csc_from_triplet")
    k = nnz.value
    return CSC(nrow, ncol, cp, ri[:k].copy(), va[:k].copy())


def csc_matvec(a: CSC, x):
    xx = _f(x)
    y = np.zeros(a.n_rows)
    _lib().ffo_csc_matvec(C.c_int(a.n_rows), C.c_int(a.n_cols), a.col_ptr.ctypes.data_as(C.c_void_p),
                          a.row_idx.ctypes.data_as(C.c_void_p), _f(a.val).ctypes.data_as(C.c_void_p),
                          xx.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))
    return y


def spmv_jvp(a: CSC, x, values_dot, x_dot):
    """spmv_jvp (src/solvers/fortfem_sparse_matrix.f90:60-91)."""
    xx, vd, xd = _f(x), _f(values_dot), _f(x_dot)
    if vd.shape[0] != a.nnz or xx.shape[0] != a.n_cols or xd.shape[0] != a.n_cols:
        raise ValueError("spmv_jvp: incompatible sizes")
    y = np.zeros(a.n_rows)
    _lib().ffo_spmv_jvp(C.c_int(a.n_rows), C.c_int(a.n_cols), a.col_ptr.ctypes.data_as(C.c_void_p),
                        a.row_idx.ctypes.data_as(C.c_void_p), _f(a.val).ctypes.data_as(C.c_void_p),
                        xx.ctypes.data_as(C.c_void_p), vd.ctypes.data_as(C.c_void_p),
                        xd.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))
    return y


def spmv_vjp(a: CSC, x, y_bar):
    """spmv_vjp (src/solvers/fortfem_sparse_matrix.f90:93-126) -> (values_bar, x_bar)."""
    xx, yb = _f(x), _f(y_bar)
    if xx.shape[0] != a.n_cols or yb.shape[0] != a.n_rows:
        raise ValueError("spmv_vjp: incompatible sizes")
    vb = np.zeros(max(a.nnz, 1))
    xb = np.zeros(a.n_cols)
    _lib().ffo_spmv_vjp(C.c_int(a.n_rows), C.c_int(a.n_cols), a.col_ptr.ctypes.data_as(C.c_void_p),
                        a.row_idx.ctypes.data_as(C.c_void_p), _f(a.val).ctypes.data_as(C.c_void_p),
                        xx.ctypes.data_as(C.c_void_p), yb.ctypes.data_as(C.c_void_p),
                        vb.ctypes.data_as(C.c_void_p), xb.ctypes.data_as(C.c_void_p))
    return vb[:a.nnz].copy(), xb


def sparse_from_dense(dense):
    d = np.asfortranarray(dense, dtype=np.float64)
    n = d.shape[0]
    flat = np.ascontiguousarray(d.reshape(-1, order="F"))
    nnz_max = int(np.count_nonzero(np.abs(d) > 1e-14))
    cp = np.zeros(n + 1, dtype=np.int32)
    ri = np.zeros(max(nnz_max, 1), dtype=np.int32)
    va = np.zeros(max(nnz_max, 1))
    nnz = C.c_int64(0)
    _lib().ffo_sparse_from_dense(C.c_int(n), flat.ctypes.data_as(C.c_void_p), cp.ctypes.data_as(C.c_void_p),
                                 ri.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p), C.byref(nnz))
    k = nnz.value
    return CSC(n, n, cp, ri[:k].copy(), va[:k].copy())


def lagrange_stiffness_csc(vertices, tets, d, q, kappa=1.0, mu=0.0, cell_kappa=None):
    """assemble_tetra_lagrange_stiffness_csc (+ per-cell kappa variant)."""
    v, t = _f(vertices), _i(tets)
    st, g, G = lagrange_dof_map(d, t)
    _check(st, "lagrange_dof_map")
    n = g.shape[1]
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(nt)
    ck = None if cell_kappa is None else _f(cell_kappa)
    _check(_lib().ffo_lagrange_stiffness_triplets(
        C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
        t.ctypes.data_as(C.c_void_p), C.c_int(d), C.c_int(q),
        None if ck is None else ck.ctypes.data_as(C.c_void_p), C.c_double(kappa), C.c_double(mu),
        g.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
        vals.ctypes.data_as(C.c_void_p)), "stiffness_triplets")
    return csc_from_triplet(G, G, rows, cols, vals), g, G


def lagrange_stiffness_csc_jvp(vertices, tets, d, q, kappa=1.0, mu=0.0, vertices_dot=None,
                               kappa_dot=0.0, mu_dot=0.0, cell_kappa=None, cell_kappa_dot=None):
    """assemble_tetra_lagrange_stiffness_csc_jvp (:336-386) -> matrix_dot CSC
    (the primal's pattern: same triplet positions)."""
    v, t = _f(vertices), _i(tets)
    vd = np.zeros_like(v) if vertices_dot is None else _f(vertices_dot)
    if vd.shape != v.shape:
        raise ValueError("vertices_dot shape differs")
    st, g, G = lagrange_dof_map(d, t)
    _check(st, "lagrange_dof_map")
    n = g.shape[1]
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(nt)
    ck = None if cell_kappa is None else _f(cell_kappa)
    ckd = None if cell_kappa_dot is None else _f(cell_kappa_dot)
    _check(_lib().ffo_lagrange_stiffness_jvp_triplets(
        C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), vd.ctypes.data_as(C.c_void_p),
        C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p), C.c_int(d), C.c_int(q),
        None if ck is None else ck.ctypes.data_as(C.c_void_p),
        None if ckd is None else ckd.ctypes.data_as(C.c_void_p), C.c_double(kappa), C.c_double(mu),
        C.c_double(kappa_dot), C.c_double(mu_dot), g.ctypes.data_as(C.c_void_p),
        rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
        vals.ctypes.data_as(C.c_void_p)), "stiffness_jvp_triplets")
    return csc_from_triplet(G, G, rows, cols, vals)


def lagrange_stiffness_csc_vjp(vertices, tets, d, q, values_bar, kappa=1.0, mu=0.0, cell_kappa=None):
    """assemble_tetra_lagrange_stiffness_csc_vjp (:388-471) ->
    (vertices_bar (nv, 3), kappa_bar, mu_bar, cell_kappa_bar (per tet))."""
    a, g, G = lagrange_stiffness_csc(vertices, tets, d, q, kappa=kappa, mu=mu, cell_kappa=cell_kappa)
    vb_in = _f(values_bar)
    if vb_in.shape[0] != a.nnz:
        raise ValueError("values_bar has the wrong size")
    v, t = _f(vertices), _i(tets)
    vb = np.zeros_like(v)
    kb = C.c_double(0)
    mb = C.c_double(0)
    ckb = np.zeros(t.shape[0])
    ck = None if cell_kappa is None else _f(cell_kappa)
    _check(_lib().ffo_lagrange_stiffness_vjp(
        C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
        t.ctypes.data_as(C.c_void_p), C.c_int(d), C.c_int(q),
        None if ck is None else ck.ctypes.data_as(C.c_void_p), C.c_double(kappa), C.c_double(mu),
        g.ctypes.data_as(C.c_void_p), a.col_ptr.ctypes.data_as(C.c_void_p),
        a.row_idx.ctypes.data_as(C.c_void_p), vb_in.ctypes.data_as(C.c_void_p),
        vb.ctypes.data_as(C.c_void_p), C.byref(kb), ckb.ctypes.data_as(C.c_void_p), C.byref(mb)),
        "stiffness_vjp")
    return vb, kb.value, mb.value, ckb


def tetra_physical_points(vertices, tets, q):
    v, t = _f(vertices), _i(tets)
    nq = _lib().ffo_tetra_duffy_count(C.c_int(q))
    pts = np.zeros((t.shape[0], nq, 3))
    _check(_lib().ffo_tetra_physical_points(C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p),
                                            C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p),
                                            C.c_int(q), pts.ctypes.data_as(C.c_void_p)), "physical_points")
    return pts


def lagrange_load_samples(vertices, tets, d, q, source_values, global_dofs, G):
    """assemble_tetra_lagrange_scalar_load_samples; source_values (T, nq)."""
    v, t, s, g = _f(vertices), _i(tets), _f(source_values), _i(global_dofs)
    rhs = np.zeros(G)
    _check(_lib().ffo_lagrange_load_samples(C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p),
                                            C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p),
                                            C.c_int(d), C.c_int(q), s.ctypes.data_as(C.c_void_p),
                                            g.ctypes.data_as(C.c_void_p), C.c_int(G),
                                            rhs.ctypes.data_as(C.c_void_p)), "load_samples")
    return rhs


def lagrange_load(vertices, tets, d, q, source, global_dofs, G):
    """assemble_tetra_lagrange_scalar_load with a vectorised source(x, y, z)."""
    pts = tetra_physical_points(vertices, tets, q)
    samples = source(pts[..., 0], pts[..., 1], pts[..., 2])
    return lagrange_load_samples(vertices, tets, d, q, samples, global_dofs, G)


def lagrange_dirichlet(vertices, tets, d, global_dofs, G, search=False):
    v, t, g = _f(vertices), _i(tets), _i(global_dofs)
    cons = np.zeros(G, dtype=np.int32)
    pts = np.zeros((G, 3))
    if search:
        _check(_lib().ffo_lagrange_dirichlet_search(C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p),
                                                    C.c_int(d), g.ctypes.data_as(C.c_void_p), C.c_int(G),
                                                    cons.ctypes.data_as(C.c_void_p)), "dirichlet_search")
    else:
        _check(_lib().ffo_lagrange_dirichlet(C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p),
                                             C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p), C.c_int(d),
                                             g.ctypes.data_as(C.c_void_p), C.c_int(G),
                                             cons.ctypes.data_as(C.c_void_p),
                                             pts.ctypes.data_as(C.c_void_p)), "dirichlet")
    return cons.astype(bool), pts


def constrained_reduction(a: CSC, constrained, cvalues, rhs):
    n = a.n_cols
    cons = _i(np.asarray(constrained, dtype=np.int32))
    cv = _f(cvalues)
    b = _f(rhs)
    free = np.zeros(n, dtype=np.int32)
    cp = np.zeros(n + 1, dtype=np.int32)
    ri = np.zeros(max(a.nnz, 1), dtype=np.int32)
    va = np.zeros(max(a.nnz, 1))
    rb = np.zeros(n)
    nf = C.c_int(0)
    nnz = C.c_int64(0)
    _check(_lib().ffo_constrained_reduction(C.c_int(n), a.col_ptr.ctypes.data_as(C.c_void_p),
                                            a.row_idx.ctypes.data_as(C.c_void_p),
                                            _f(a.val).ctypes.data_as(C.c_void_p),
                                            cons.ctypes.data_as(C.c_void_p), cv.ctypes.data_as(C.c_void_p),
                                            b.ctypes.data_as(C.c_void_p), C.byref(nf),
                                            free.ctypes.data_as(C.c_void_p), cp.ctypes.data_as(C.c_void_p),
                                            ri.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                            C.byref(nnz), rb.ctypes.data_as(C.c_void_p)), "reduction")
    f = nf.value
    k = nnz.value
    return CSC(f, f, cp[: f + 1].copy(), ri[:k].copy(), va[:k].copy()), free[:f].copy(), rb[:f].copy()


# --------------------------------------------------------------- Krylov ---
class _Stats(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("final_residual", C.c_double)]


@dataclass
class SolverStats:
    converged: bool
    iterations: int
    final_residual: float


METHODS = {"cg": 0, "pcg": 1, "auto": 1}
PRECONDS = {"none": 0, "jacobi": 1}
TOL_TYPES = {"relative": 0, "absolute": 1}
BREAKDOWN = {"absolute": 0, "relative": 1}


def solve_sparse(a: CSC, b, x0=None, method="pcg", preconditioner="jacobi", tolerance=1e-6,
                 tolerance_type="relative", max_iterations=1000, breakdown="absolute"):
    """solve_sparse restated (advanced_solvers.f90:135-195)."""
    n = a.n_cols
    x = np.zeros(n) if x0 is None else _f(x0).copy()
    bb = _f(b)
    st = _Stats()
    _lib().ffo_solve_sparse(C.c_int(n), a.col_ptr.ctypes.data_as(C.c_void_p),
                            a.row_idx.ctypes.data_as(C.c_void_p), _f(a.val).ctypes.data_as(C.c_void_p),
                            bb.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                            C.c_int(METHODS[method]), C.c_int(PRECONDS[preconditioner]),
                            C.c_double(tolerance), C.c_int(TOL_TYPES[tolerance_type]),
                            C.c_int(max_iterations), C.c_int(BREAKDOWN[breakdown]), C.byref(st))
    return x, SolverStats(bool(st.converged), int(st.iterations), float(st.final_residual))


def solve_sparse_jvp(a: CSC, x, values_dot, b_dot, **opts):
    """Sparse form of pcg_solve_jvp / cg_solve_jvp (advanced_solvers.f90:387-419,
    458-489): x_dot solves A x_dot = b_dot - A_dot x from x_dot = 0 with the
    same solver; A_dot is the CSC value tangent on A's pattern.
    Returns (x_dot, stats, status) with status 0 / 2 (not converged)."""
    a_dot = CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, _f(values_dot))
    rhs = _f(b_dot) - csc_matvec(a_dot, x)
    xd, st = solve_sparse(a, rhs, np.zeros(a.n_cols), **opts)
    return xd, st, 0 if st.converged else 2


def solve_sparse_vjp(a: CSC, x, x_bar, **opts):
    """Sparse form of pcg_solve_vjp / cg_solve_vjp (advanced_solvers.f90:421-456,
    491-522): lambda solves A^T lambda = x_bar; b_bar = lambda and A_bar =
    -lambda x^T, here kept on A's stored entries.
    Returns (values_bar, b_bar, stats, status)."""
    at = sparse_from_dense(a.to_dense().T) if not _is_symmetric(a) else a
    lam, st = solve_sparse(at, _f(x_bar), np.zeros(a.n_cols), **opts)
    if not st.converged:
        return np.zeros(a.nnz), np.zeros(a.n_cols), st, 2
    cols = np.repeat(np.arange(a.n_cols), np.diff(a.col_ptr))
    vb = -lam[a.row_idx - 1] * _f(x)[cols]
    return vb, lam, st, 0


def _is_symmetric(a: CSC):
    d = a.to_dense()
    return bool(np.array_equal(d, d.T))


# ------------------------------------------------------------------ 2D ---
def p1_2d_dense(vertices, tris, mass=0.0, stiffness=1.0):
    v, t = _f(vertices), _i(tris)
    n = v.shape[0]
    m = np.zeros((n, n))  # C-order m[c, r] == matrix(r, c)
    _check(_lib().ffo_p1_2d_assemble_dense(C.c_int(n), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
                                           t.ctypes.data_as(C.c_void_p), C.c_double(mass),
                                           C.c_double(stiffness), m.ctypes.data_as(C.c_void_p)), "p1_2d")
    return np.asfortranarray(m.T)


def p1_2d_load(vertices, tris, f=1.0):
    v, t = _f(vertices), _i(tris)
    b = np.zeros(v.shape[0])
    _check(_lib().ffo_p1_2d_load(C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
                                 t.ctypes.data_as(C.c_void_p), C.c_double(f), b.ctypes.data_as(C.c_void_p)),
           "p1_2d_load")
    return b


def p1_2d_symmetric_dirichlet(dense, rhs, is_boundary, value=0.0):
    n = dense.shape[0]
    m = np.ascontiguousarray(np.asfortranarray(dense).reshape(-1, order="F"))
    b = _f(rhs).copy()
    bnd = _i(np.asarray(is_boundary, dtype=np.int32))
    _lib().ffo_p1_2d_symmetric_dirichlet(C.c_int(n), m.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                         bnd.ctypes.data_as(C.c_void_p), C.c_double(value))
    return m.reshape((n, n), order="F"), b


def p1_2d_poisson_system(n=65, f=1.0, g=0.0):
    """Configuration 1: unit_square_mesh(n), f, symmetric elimination, then
    sparse_from_dense -- the CSC the reference's Jacobi-PCG sees."""
    v, t, bnd = unit_square_mesh(n)
    a = p1_2d_dense(v, t, 0.0, 1.0)
    b = p1_2d_load(v, t, f)
    a, b = p1_2d_symmetric_dirichlet(a, b, bnd, g)
    return sparse_from_dense(a), b, (v, t, bnd)


# ------------------------------------------------------------- Nedelec ---
def nedelec_dof_count(k):
    return _lib().ffo_nedelec_dof_count(C.c_int(k))


def nedelec_coefficients(k):
    n = nedelec_dof_count(k)
    c = np.zeros((n, n))
    _check(_lib().ffo_nedelec_coefficients(C.c_int(k), c.ctypes.data_as(C.c_void_p)), "nedelec_coefficients")
    return c


def nedelec_eval(k, point, coeff=None):
    """reference basis values/curls (n, 3); k = 1 gives the Whitney functions"""
    n = nedelec_dof_count(k)
    p = _f(point)
    v = np.zeros((n, 3))
    c = np.zeros((n, 3))
    if k == 1:
        _lib().ffo_whitney(p.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
    else:
        cf = nedelec_coefficients(k) if coeff is None else _f(coeff)
        _lib().ffo_nedelec_eval(C.c_int(k), cf.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p),
                                v.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
    return v, c


def face_basis_to_local(k, perm):
    s = k * (k - 1)
    T = np.zeros((s, s))
    pp = _i(perm)
    _check(_lib().ffo_face_basis_to_local(C.c_int(k), pp.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p)),
           "face_basis_to_local")
    return T


def edge_dof_map(tets, search=False):
    t = _i(tets)
    T = t.shape[0]
    edges = np.zeros((6 * T, 2), dtype=np.int32)
    gd = np.zeros((T, 6), dtype=np.int32)
    orient = np.zeros((T, 6), dtype=np.int32)
    ne = C.c_int(0)
    fn = _lib().ffo_edge_dof_map_search if search else _lib().ffo_edge_dof_map
    st = fn(C.c_int(T), t.ctypes.data_as(C.c_void_p), edges.ctypes.data_as(C.c_void_p), gd.ctypes.data_as(C.c_void_p),
            orient.ctypes.data_as(C.c_void_p), C.byref(ne))
    return st, edges[:ne.value].copy(), gd, orient


def nedelec_dof_map(k, tets):
    t = _i(tets)
    T = t.shape[0]
    n = nedelec_dof_count(k)
    gd = np.zeros((T, n), dtype=np.int32)
    eo = np.zeros((T, 6), dtype=np.int32)
    fp = np.zeros((T, 4, 3), dtype=np.int32)
    nd, ne, nf = C.c_int(0), C.c_int(0), C.c_int(0)
    st = _lib().ffo_nedelec_dof_map(C.c_int(k), C.c_int(T), t.ctypes.data_as(C.c_void_p), gd.ctypes.data_as(C.c_void_p),
                                    eo.ctypes.data_as(C.c_void_p), fp.ctypes.data_as(C.c_void_p), C.byref(nd),
                                    C.byref(ne), C.byref(nf))
    _check(st, "nedelec_dof_map")
    return gd, eo, fp, nd.value, ne.value, nf.value


def nedelec_basis_transform(k, edge_orient, face_perm):
    n = nedelec_dof_count(k)
    T = np.zeros((n, n))
    eo = _i(edge_orient)
    fp = _i(face_perm)
    _check(_lib().ffo_nedelec_basis_transform(C.c_int(k), eo.ctypes.data_as(C.c_void_p), fp.ctypes.data_as(C.c_void_p),
                                              T.ctypes.data_as(C.c_void_p)), "basis_transform")
    return T


def nedelec_element(k, vertices4, qdeg=None, alpha=1.0, beta=1.0, tensor=None):
    n = nedelec_dof_count(k)
    cf = nedelec_coefficients(k) if k >= 2 else np.zeros(1)
    if qdeg is None:
        qdeg = 2 * k + 2 if tensor is not None else 2 * k
    m = np.zeros((n, n))  # C order m[c, r]
    K = None if tensor is None else np.ascontiguousarray(np.asarray(tensor, dtype=np.float64).T)  # column-major
    st = _lib().ffo_nedelec_element(C.c_int(k), cf.ctypes.data_as(C.c_void_p), _f(vertices4).ctypes.data_as(C.c_void_p),
                                    C.c_int(qdeg), C.c_double(alpha), None if K is None else K.ctypes.data_as(C.c_void_p),
                                    C.c_double(beta), m.ctypes.data_as(C.c_void_p))
    return st, m.T.copy()


def nedelec_rule_points(k, qdeg=None):
    """Points of the Nedelec element rule (the 4-point rule at k = 1, Duffy of
    degree qdeg, default 2k + 2, otherwise): (nq, 4) rows x, y, z, w."""
    if qdeg is None:
        qdeg = 2 * k + 2
    cap = 4 if k == 1 else int(_lib().ffo_tetra_duffy_count(C.c_int(qdeg)))
    out = np.zeros(4 * cap)
    n = C.c_int(0)
    _check(_lib().ffo_nedelec_rule_points(C.c_int(k), C.c_int(qdeg), C.c_int(cap),
                                          out.ctypes.data_as(C.c_void_p), C.byref(n)), "rule_points")
    return out[:4 * n.value].reshape(n.value, 4)


def nedelec_element_points(k, vertices4, point_tensors, beta=1.0, qdeg=None):
    """The weighted element (:2406-2514) with a tensor at every point of the
    rule: point_tensors (nq, 3, 3) in the order of nedelec_rule_points."""
    n = nedelec_dof_count(k)
    cf = nedelec_coefficients(k) if k >= 2 else np.zeros(1)
    if qdeg is None:
        qdeg = 2 * k + 2
    K = np.ascontiguousarray(np.transpose(np.asarray(point_tensors, dtype=np.float64), (0, 2, 1)).reshape(-1))
    m = np.zeros((n, n))
    st = _lib().ffo_nedelec_element_points(C.c_int(k), cf.ctypes.data_as(C.c_void_p),
                                           _f(vertices4).ctypes.data_as(C.c_void_p), C.c_int(qdeg),
                                           K.ctypes.data_as(C.c_void_p), C.c_double(beta),
                                           m.ctypes.data_as(C.c_void_p))
    return st, m.T.copy()


def nedelec_csc_points(k, vertices, tets, point_tensors, beta=1.0):
    """assemble_tetra_nedelec_weighted_csc restated with the coefficient
    evaluated at every point of the weighted rule: point_tensors (T, nq, 3, 3),
    points nedelec_rule_points(k) mapped into each cell as v1 + J xi."""
    v, t = _f(vertices), _i(tets)
    n = nedelec_dof_count(k)
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(nt)
    nd = C.c_int(0)
    K = np.ascontiguousarray(np.transpose(np.asarray(point_tensors, dtype=np.float64), (0, 1, 3, 2)).reshape(-1))
    _check(_lib().ffo_nedelec_triplets_points(C.c_int(k), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p),
                                              C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p), C.c_double(beta),
                                              K.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p),
                                              cols.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p),
                                              C.byref(nd)), "nedelec_triplets_points")
    return csc_from_triplet(nd.value, nd.value, rows, cols, vals)


def mapped_points(vertices, tets, ref_points):
    """x = v1 + J xi for every cell and reference point: (T, nq, 3)."""
    v, t = np.asarray(vertices, dtype=np.float64), np.asarray(tets) - 1
    v0 = v[t[:, 0]]
    J = np.stack([v[t[:, 1]] - v0, v[t[:, 2]] - v0, v[t[:, 3]] - v0], axis=2)  # (T, 3, 3) columns
    xi = np.asarray(ref_points)[:, :3]
    return v0[:, None, :] + np.einsum("tij,qj->tqi", J, xi)


def nedelec_csc(k, vertices, tets, alpha=1.0, beta=1.0, cell_tensor=None):
    """assemble_tetra_nedelec_curl_mass_csc / _weighted_csc restated.
    cell_tensor: (T, 3, 3) per-cell tensors (weighted variant)."""
    v, t = _f(vertices), _i(tets)
    n = nedelec_dof_count(k)
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(nt)
    nd = C.c_int(0)
    K = None
    if cell_tensor is not None:
        K = np.ascontiguousarray(np.transpose(np.asarray(cell_tensor, dtype=np.float64), (0, 2, 1)).reshape(-1))
    _check(_lib().ffo_nedelec_triplets(C.c_int(k), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p),
                                       C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p), C.c_double(alpha),
                                       C.c_double(beta), None if K is None else K.ctypes.data_as(C.c_void_p),
                                       rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                                       vals.ctypes.data_as(C.c_void_p), C.byref(nd)), "nedelec_triplets")
    return csc_from_triplet(nd.value, nd.value, rows, cols, vals)


def nedelec_csc_jvp(k, vertices, tets, alpha=1.0, beta=1.0, vertices_dot=None, alpha_dot=0.0,
                    beta_dot=0.0, cell_alpha=None, cell_alpha_dot=None):
    """assemble_tetra_nedelec_curl_mass_csc_jvp (:2114-2180) -> matrix_dot CSC."""
    v, t = _f(vertices), _i(tets)
    vd = np.zeros_like(v) if vertices_dot is None else _f(vertices_dot)
    n = nedelec_dof_count(k)
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(nt)
    nd = C.c_int(0)
    ca = None if cell_alpha is None else _f(cell_alpha)
    cad = None if cell_alpha_dot is None else _f(cell_alpha_dot)
    _check(_lib().ffo_nedelec_jvp_triplets(
        C.c_int(k), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), vd.ctypes.data_as(C.c_void_p),
        C.c_int(t.shape[0]), t.ctypes.data_as(C.c_void_p), C.c_double(alpha), C.c_double(beta),
        C.c_double(alpha_dot), C.c_double(beta_dot), None if ca is None else ca.ctypes.data_as(C.c_void_p),
        None if cad is None else cad.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p),
        cols.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p), C.byref(nd)), "nedelec_jvp")
    return csc_from_triplet(nd.value, nd.value, rows, cols, vals)


def nedelec_csc_vjp(k, vertices, tets, values_bar, alpha=1.0, beta=1.0, cell_alpha=None):
    """assemble_tetra_nedelec_curl_mass_csc_vjp (:2182-2287) ->
    (vertices_bar (nv, 3), alpha_bar, beta_bar, cell_alpha_bar)."""
    a = nedelec_csc_jvp(k, vertices, tets)  # the primal pattern
    vb_in = _f(values_bar)
    if vb_in.shape[0] != a.nnz:
        raise ValueError("values_bar has the wrong size")
    v, t = _f(vertices), _i(tets)
    vb = np.zeros_like(v)
    ab = C.c_double(0)
    bb = C.c_double(0)
    cab = np.zeros(t.shape[0])
    ca = None if cell_alpha is None else _f(cell_alpha)
    _check(_lib().ffo_nedelec_vjp(
        C.c_int(k), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
        t.ctypes.data_as(C.c_void_p), C.c_double(alpha), C.c_double(beta),
        None if ca is None else ca.ctypes.data_as(C.c_void_p), a.col_ptr.ctypes.data_as(C.c_void_p),
        a.row_idx.ctypes.data_as(C.c_void_p), vb_in.ctypes.data_as(C.c_void_p),
        vb.ctypes.data_as(C.c_void_p), C.byref(ab), cab.ctypes.data_as(C.c_void_p), C.byref(bb)),
        "nedelec_vjp")
    return vb, ab.value, bb.value, cab


def nedelec_pml_csc(k, vertices, tets, stretch, wave_number, kind=0):
    """assemble_tetra_nedelec_pml_csc (kind 0, stretch (T, 3) complex) and
    assemble_tetra_nedelec_curvilinear_pml_csc (kind 1, stretch (T, 3, 3)
    complex, S[t][i, j] = stretch(i, j, t)) restated -> CSC with complex values."""
    v, t = _f(vertices), _i(tets)
    S = np.asarray(stretch, dtype=np.complex128)
    if kind == 1:
        S = np.transpose(S, (0, 2, 1))  # column-major 3 x 3 per cell
    st = np.ascontiguousarray(np.stack([S.real, S.imag], axis=-1).reshape(-1))
    n = nedelec_dof_count(k)
    nt = t.shape[0] * n * n
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    vals = np.zeros(2 * nt)
    nd = C.c_int(0)
    _check(_lib().ffo_nedelec_pml_triplets(
        C.c_int(k), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
        t.ctypes.data_as(C.c_void_p), C.c_int(kind), st.ctypes.data_as(C.c_void_p), C.c_double(wave_number),
        rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p),
        C.byref(nd)), "nedelec_pml_triplets")
    re = csc_from_triplet(nd.value, nd.value, rows, cols, vals[0::2])
    im = csc_from_triplet(nd.value, nd.value, rows, cols, vals[1::2])
    return CSC(re.n_rows, re.n_cols, re.col_ptr, re.row_idx, re.val + 1j * im.val)


def nedelec_load_rule(k):
    """The load quadrature: tetra_degree_two_quadrature at k = 1
    (src/assembly/assembly_tetra_nedelec_3d.f90:2570-2581), the Duffy rule
    of degree 2k + 2 otherwise (:546-547).  -> (points (nq, 3), weights)."""
    if k == 1:
        a = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
        b = (5.0 - np.sqrt(5.0)) / 20.0
        return np.array([[b, b, b], [a, b, b], [b, a, b], [b, b, a]]), np.full(4, 1.0 / 24.0)
    x, y, z, w = tetra_duffy(2 * k + 2)
    return np.stack([x, y, z], axis=1), np.asarray(w)


def nedelec_vector_load(k, vertices, tets, source):
    """assemble_tetra_nedelec_vector_load (k = 1: :2002-2040 with the element
    :2516-2553) and assemble_tetra_nedelec_vector_load_order (k >= 2: :71-127
    with :525-582) restated: b_r = sum_q det w_q f(x_q) . J^-T v_r(xi_q),
    x_q = v1 + J xi_q, oriented by the edge signs (k = 1) or T^T (k >= 2) and
    summed into the global vector.  source(x, y, z) -> 3-vector.  Python
    loops: small meshes only."""
    v, t = _f(vertices), _i(tets)
    n = nedelec_dof_count(k)
    pts, w = nedelec_load_rule(k)
    cf = nedelec_coefficients(k) if k >= 2 else None
    V = [nedelec_eval(k, p, cf)[0] for p in pts]  # (n, 3) per point
    if k == 1:
        st, edges, gd, orient = edge_dof_map(t)
        _check(st, "edge_dof_map")
        nd = edges.shape[0]
    else:
        gd, eo, fp, nd, ne, nf = nedelec_dof_map(k, t)
    vec = np.zeros(nd)
    for e in range(t.shape[0]):
        x = v[t[e] - 1]
        J = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
        det = np.linalg.det(J)
        Jinv = np.linalg.inv(J)
        elem = np.zeros(n)
        for q in range(len(w)):
            phys = V[q] @ Jinv  # row j: (J^-T v_j)^T
            f = np.asarray(source(*(x[0] + J @ pts[q])), dtype=np.float64)
            elem += det * w[q] * (phys @ f)
        if k == 1:
            vec[gd[e] - 1] += orient[e] * elem
        else:
            T = nedelec_basis_transform(k, eo[e], fp[e])
            vec[gd[e] - 1] += T.T @ elem
    return vec


def nedelec_field(k, vertices, tets, dofs, cell, xi):
    """The global field sum_i dofs_i psi_i at reference point xi of a cell:
    the cell's local coefficients T d_loc on the mapped reference basis."""
    v, t = _f(vertices), _i(tets)
    x = v[t[cell] - 1]
    J = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
    Jinv = np.linalg.inv(J)
    cf = nedelec_coefficients(k) if k >= 2 else None
    vals = nedelec_eval(k, xi, cf)[0] @ Jinv
    if k == 1:
        st, edges, gd, orient = edge_dof_map(t)
        c = orient[cell] * dofs[gd[cell] - 1]
    else:
        gd, eo, fp, nd, ne, nf = nedelec_dof_map(k, t)
        c = nedelec_basis_transform(k, eo[cell], fp[cell]) @ dofs[gd[cell] - 1]
    return c @ vals


# ------------------------------------------- threaded CPU-baseline stages ---
def par_lagrange_stiffness_csc(vertices, tets, d, q, global_dofs, G, nthreads, kappa=1.0, mu=0.0,
                               cell_kappa=None):
    """lagrange_stiffness_csc's triplet stage and csc_from_triplet on
    nthreads OpenMP threads, given the dof map: the same CSC, for timing the
    CPU baseline on all host cores."""
    v, t, g = _f(vertices), _i(tets), _i(global_dofs)
    n = g.shape[1]
    nt = t.shape[0] * n * n
    rows = np.empty(nt, dtype=np.int32)
    cols = np.empty(nt, dtype=np.int32)
    vals = np.empty(nt)
    ck = None if cell_kappa is None else _f(cell_kappa)
    _check(_lib().ffo_par_lagrange_stiffness_triplets(
        C.c_int(nthreads), C.c_int(v.shape[0]), v.ctypes.data_as(C.c_void_p), C.c_int(t.shape[0]),
        t.ctypes.data_as(C.c_void_p), C.c_int(d), C.c_int(q),
        None if ck is None else ck.ctypes.data_as(C.c_void_p), C.c_double(kappa), C.c_double(mu),
        g.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
        vals.ctypes.data_as(C.c_void_p)), "par_stiffness_triplets")
    cp = np.zeros(G + 1, dtype=np.int32)
    ri = np.empty(max(nt, 1), dtype=np.int32)
    va = np.empty(max(nt, 1))
    nnz = C.c_int64(0)
    _check(_lib().ffo_par_csc_from_triplet(C.c_int(nthreads), C.c_int(G), C.c_int(G), C.c_int64(nt),
                                           rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                                           vals.ctypes.data_as(C.c_void_p), cp.ctypes.data_as(C.c_void_p),
                                           ri.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                           C.byref(nnz)), "par_csc_from_triplet")
    k = nnz.value
    del rows, cols, vals
    return CSC(G, G, cp, ri[:k].copy(), va[:k].copy())


def par_solve_sparse(a: CSC, b, nthreads, x0=None, method="pcg", preconditioner="jacobi", tolerance=1e-6,
                     tolerance_type="relative", max_iterations=1000, breakdown="absolute"):
    """solve_sparse's loop on nthreads threads (symmetric CSC read as CSR)."""
    n = a.n_cols
    x = np.zeros(n) if x0 is None else _f(x0).copy()
    bb = _f(b)
    st = _Stats()
    _check(_lib().ffo_par_solve_sparse(C.c_int(nthreads), C.c_int(n), a.col_ptr.ctypes.data_as(C.c_void_p),
                                       a.row_idx.ctypes.data_as(C.c_void_p), _f(a.val).ctypes.data_as(C.c_void_p),
                                       bb.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                       C.c_int(METHODS[method]), C.c_int(PRECONDS[preconditioner]),
                                       C.c_double(tolerance), C.c_int(TOL_TYPES[tolerance_type]),
                                       C.c_int(max_iterations), C.c_int(BREAKDOWN[breakdown]), C.byref(st)),
           "par_solve_sparse")
    return x, SolverStats(bool(st.converged), int(st.iterations), float(st.final_residual))
