"""fortfem_amd -- MI355X (gfx950) drop-in for fortfem's sparse-assembly +
Jacobi-PCG path.

The compute lives in ``libfortfem_hip.so`` (hand-written HIP kernels behind the
C-ABI of ``include/fortfem_hip.h``).  This module is the Python host side: it
mirrors the reference's procedure names, argument meaning and status
behaviour so that callers (and the parity tests) read like fortfem code:

=====================================  ==============================================
reference (lazy-fortran/fortfem)        here
=====================================  ==============================================
generate_structured_tetra_box_mesh      :func:`generate_structured_tetra_box_mesh`
  (src/mesh/structured_tetra_box_mesh.f90:12-70)
unit_square_mesh (fortfem_api_mesh.f90)  :func:`unit_square_mesh`
assemble_tetra_lagrange_stiffness_csc   :func:`assemble_tetra_lagrange_stiffness_csc`
  (assembly_tetra_lagrange_arbitrary_order_3d.f90:276-334)
assemble_tetra_lagrange_scalar_load     :func:`assemble_tetra_lagrange_scalar_load`
  (:705-774) / _load_samples (:473-528)  :func:`assemble_tetra_lagrange_scalar_load_samples`
solver_options / solver_options_t       :func:`solver_options`
  (advanced_solvers.f90:32-44,1430-1465)
solve_sparse (advanced_solvers.f90:135)  :func:`solve_sparse`
spmv (fortfem_sparse_matrix.f90:46-58)   :func:`spmv`
solve_tetra_lagrange_poisson            :func:`solve_tetra_lagrange_poisson`
  (fortfem_tetra_lagrange_solver_3d.f90:50-129)
=====================================  ==============================================

Failures are reported like the reference: assembly routines return a
``FortsparseStatus`` whose ``code`` is nonzero on failure (the reference presets
FORTSPARSE_INVALID_MATRIX and clears it on success), and programmer errors
(``error stop`` in the reference, e.g. size mismatches in ``spmv`` and
``solve_sparse``) raise.  A missing native library or GPU raises: there is no
CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from ._capi import FFHError

FORTSPARSE_INVALID_MATRIX = 2  # status preset of the reference assemblers


@dataclass
class FortsparseStatus:
    code: int = 0
    msg: str = ""


@dataclass
class CSC:
    """fortsparse csc_t: 1-based col_ptr(n+1), row_idx(nnz), val(nnz)."""

    nrow: int
    ncol: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_idx.shape[0])

    def to_dense(self) -> np.ndarray:
        a = np.zeros((self.nrow, self.ncol))
        cols = np.repeat(np.arange(self.ncol), np.diff(self.col_ptr))
        np.add.at(a, (self.row_idx - 1, cols), self.val)
        return a


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


_OWNED = weakref.WeakSet()  # problems, matrices and servers, closed before their context


@atexit.register
def _shutdown():
    for obj in list(_OWNED):
        try:
            obj.close()
        except Exception:
            pass
    if _DEFAULT is not None:
        _DEFAULT.close()


class Context:
    """One device + stream (ffh_context)."""

    def __init__(self, device: int = 0):
        self._lib = capi.load()
        h = C.c_void_p()
        capi.check(self._lib.ffh_init(device, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            for obj in list(_OWNED):
                if getattr(obj, "ctx", None) is self:
                    obj.close()
            self._lib.ffh_finalize(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code):
        capi.check(code, self.handle)

    def set_profiling(self, enable: bool):
        self.check(self._lib.ffh_set_profiling(self.handle, 1 if enable else 0))

    def last_solve_profile(self, which: int = 0):
        cnt = C.c_int(0)
        ms = C.c_double(0.0)
        self.check(self._lib.ffh_last_solve_profile(self.handle, which, C.byref(cnt), C.byref(ms)))
        return cnt.value, ms.value


_DEFAULT: Context | None = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


# ------------------------------------------------------------------- options
@dataclass
class SolverOptionsT:
    """solver_options_t (advanced_solvers.f90:32-44) plus this library's fields."""

    method: str = "auto"
    tolerance: float = 1.0e-6
    max_iterations: int = 1000
    preconditioner: str = "none"
    tolerance_type: str = "relative"
    precision: str = "fp64"          # "fp64" | "mixed"
    breakdown: str = "absolute"      # "absolute" (reference) | "relative"
    inner_tolerance: float = 1.0e-4
    max_refinements: int = 50
    check_interval: int = 0
    use_graph: bool = False

    def to_c(self) -> capi.SolverOptions:
        o = capi.SolverOptions()
        capi.load().ffh_solver_options_default(C.byref(o))
        m = self.method.strip().lower()
        if m not in ("auto", "cg", "pcg"):
            m = "pcg"  # solve_sparse default branch (:184-187)
        o.method = 0 if m == "cg" else 1
        o.preconditioner = 1 if self.preconditioner.strip().lower() == "jacobi" else 0
        o.tolerance_type = 1 if self.tolerance_type.strip().lower() == "absolute" else 0
        o.max_iterations = int(self.max_iterations)
        o.tolerance = float(self.tolerance)
        o.precision = capi.MIXED if self.precision == "mixed" else capi.FP64
        o.breakdown_mode = 1 if self.breakdown == "relative" else 0
        o.inner_tolerance = float(self.inner_tolerance)
        o.max_refinements = int(self.max_refinements)
        o.check_interval = int(self.check_interval)
        o.use_graph = 1 if self.use_graph else 0
        return o


def solver_options(method=None, tolerance=None, max_iterations=None, preconditioner=None,
                   tolerance_type=None, **extra) -> SolverOptionsT:
    """solver_options() constructor (advanced_solvers.f90:1430-1465)."""
    o = SolverOptionsT()
    if method is not None:
        o.method = method
    if tolerance is not None:
        o.tolerance = tolerance
    if max_iterations is not None:
        o.max_iterations = max_iterations
    if preconditioner is not None:
        o.preconditioner = preconditioner
    if tolerance_type is not None:
        o.tolerance_type = tolerance_type
    for k, v in extra.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown solver option {k}")
        setattr(o, k, v)
    return o


@dataclass
class SolverStatsT:
    """solver_stats_t (advanced_solvers.f90:47-57) plus this library's fields."""

    converged: bool = False
    iterations: int = 0
    final_residual: float = 0.0
    solve_time: float = 0.0
    method_used: str = ""
    refinements: int = 0
    true_residual: float = 0.0
    initial_residual: float = 0.0

    @classmethod
    def from_c(cls, s: capi.SolverStats, method: str):
        return cls(bool(s.converged), int(s.iterations), float(s.final_residual),
                   float(s.solve_time), method, int(s.refinements), float(s.true_residual),
                   float(s.initial_residual))


# -------------------------------------------------------------------- meshes
def generate_structured_tetra_box_mesh(bounds, counts):
    """bounds = [[xlo, ylo, zlo], [xhi, yhi, zhi]] (bounds(3,2)); returns
    (vertices (nv,3), tetrahedra (T,4) 1-based, status)."""
    lib = capi.load()
    b = _f64(np.asarray(bounds, dtype=np.float64).reshape(2, 3).reshape(-1))
    c = _i32(np.asarray(counts).reshape(3))
    nv = C.c_int(0)
    nt = C.c_int(0)
    st = lib.ffh_box_mesh_size(_ptr(c), C.byref(nv), C.byref(nt))
    if st != 0:
        return np.zeros((0, 3)), np.zeros((0, 4), dtype=np.int32), 1
    v = np.zeros((nv.value, 3))
    t = np.zeros((nt.value, 4), dtype=np.int32)
    st = lib.ffh_box_mesh(_ptr(b), _ptr(c), _ptr(v), _ptr(t))
    return v, t, (0 if st == 0 else 1)


def unit_box_mesh(n):
    v, t, st = generate_structured_tetra_box_mesh([[0, 0, 0], [1, 1, 1]], [n, n, n])
    if st:
        raise ValueError("invalid box mesh")
    return v, t


def unit_square_mesh(n):
    """unit_square_mesh(n): (vertices (n^2,2), triangles (2(n-1)^2,3) 1-based, is_boundary)."""
    lib = capi.load()
    v = np.zeros((n * n, 2))
    t = np.zeros((2 * (n - 1) ** 2, 3), dtype=np.int32)
    b = np.zeros(n * n, dtype=np.int32)
    capi.check(lib.ffh_unit_square_mesh(n, _ptr(v), _ptr(t), _ptr(b)))
    return v, t, b.astype(bool)


# ------------------------------------------------------------------- problem
class Problem:
    """A mesh + discrete space + operator resident in HBM (ffh_problem)."""

    def __init__(self, vertices, cells, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._lib = capi.load()
        v = _f64(vertices)
        c = _i32(cells)
        self.dim = v.shape[1]
        h = C.c_void_p()
        self.ctx.check(self._lib.ffh_problem_create(self.ctx.handle, self.dim, v.shape[0], _ptr(v),
                                                    c.shape[0], _ptr(c), C.byref(h)))
        self.handle = h
        self.ncell = c.shape[0]
        self.nvert = v.shape[0]
        _OWNED.add(self)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.ffh_problem_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code):
        self.ctx.check(code)

    def pattern_build(self, family: int, order: int):
        n = C.c_int(0)
        nnz = C.c_int64(0)
        self._check(self._lib.ffh_pattern_build(self.handle, family, order, C.byref(n), C.byref(nnz)))
        self._family, self._order = family, order
        return n.value, nnz.value

    def sizes(self):
        vals = [C.c_int(0), C.c_int64(0), C.c_int(0), C.c_int64(0), C.c_int(0), C.c_int(0)]
        self._check(self._lib.ffh_problem_sizes(self.handle, *(C.byref(x) for x in vals)))
        keys = ("n_global", "nnz", "n_free", "nnz_free", "n_local", "n_colors")
        return {k: x.value for k, x in zip(keys, vals)}

    def global_dofs(self, base=1):
        s = self.sizes()
        g = np.zeros((self.ncell, s["n_local"]), dtype=np.int32)
        self._check(self._lib.ffh_global_dofs_download(self.handle, base, _ptr(g)))
        return g

    def colouring(self, base=0):
        """(color_start, cells): the element colouring of the pattern build,
        cells listed colour by colour (ffh_problem_colouring)."""
        s = self.sizes()
        cs = np.zeros(s["n_colors"] + 1, dtype=np.int32)
        cells = np.zeros(self.ncell, dtype=np.int32)
        self._check(self._lib.ffh_problem_colouring(self.handle, base, _ptr(cs), _ptr(cells)))
        return cs, cells

    def assemble(self, quadrature_degree, kappa=1.0, mass=0.0, cell_kappa=None, precision="fp64",
                 reuse_coefficients=False, cell_tensor=None, reuse_tensor=False, point_tensor=None,
                 reuse_point_tensor=False):
        """Assemble the family's operator.  H1: kappa (stiffness) and mass;
        Nedelec: kappa = curl coefficient alpha, mass = beta, or cell_tensor
        (T, 3, 3) symmetric tensors on the curl term (weighted variant), or
        point_tensor (T, nq, 3, 3): the tensor at every point of
        nedelec_load_points(order) mapped into each cell (the reference's
        coefficient(x_q)); reuse_point_tensor keeps the last uploaded ones."""
        prec = capi.FP32 if precision == "fp32" else capi.FP64
        if reuse_point_tensor:
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_POINT_TENSOR,
                                               None, float(mass), prec))
        elif point_tensor is not None:
            K = np.asarray(point_tensor, dtype=np.float64)
            if getattr(self, "_family", None) != capi.FAMILY_NEDELEC:
                raise ValueError("point tensors need a Nedelec pattern")
            nq = nedelec_load_points(self._order).shape[0]
            if K.shape != (self.ncell, nq, 3, 3):
                raise ValueError(f"point_tensor must have shape ({self.ncell}, {nq}, 3, 3)")
            K = np.ascontiguousarray(K.transpose(0, 1, 3, 2))
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_POINT_TENSOR,
                                               _ptr(K.reshape(-1)), float(mass), prec))
        elif reuse_tensor:
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_CELL_TENSOR,
                                               None, float(mass), prec))
        elif cell_tensor is not None:
            K = np.ascontiguousarray(np.asarray(cell_tensor, dtype=np.float64).transpose(0, 2, 1))
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_CELL_TENSOR,
                                               _ptr(K.reshape(-1)), float(mass), prec))
        elif cell_kappa is not None or reuse_coefficients:
            ck = None if reuse_coefficients else _f64(cell_kappa)
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_CELL,
                                               _ptr(ck), float(mass), prec))
        else:
            k = _f64([kappa])
            self._check(self._lib.ffh_assemble(self.handle, quadrature_degree, capi.COEFF_SCALAR,
                                               _ptr(k), float(mass), prec))

    def _coefficients(self, kappa, cell_kappa):
        if cell_kappa is not None:
            ck = _f64(cell_kappa)
            if ck.shape[0] != self.ncell:
                raise ValueError("cell_kappa needs one value per cell")
            return capi.COEFF_CELL, ck
        return capi.COEFF_SCALAR, _f64([kappa])

    def assemble_jvp(self, quadrature_degree, kappa=1.0, mass=0.0, cell_kappa=None, vertices_dot=None,
                     kappa_dot=0.0, cell_kappa_dot=None, mass_dot=0.0):
        """Tangent of the H1 or Nedelec matrix values (ffh_assemble_jvp) on
        the pattern of matrix(FULL); vertices_dot is (nv, 3).  Nedelec: kappa
        is the curl coefficient, mass the mass coefficient."""
        kind, k = self._coefficients(kappa, cell_kappa)
        if kind == capi.COEFF_CELL:
            kd = np.zeros(self.ncell) if cell_kappa_dot is None else _f64(cell_kappa_dot)
            if kd.shape[0] != self.ncell:
                raise ValueError("cell_kappa_dot needs one value per cell")
        else:
            if cell_kappa_dot is not None:
                raise ValueError("cell_kappa_dot needs cell_kappa")
            kd = _f64([kappa_dot])
        vd = None if vertices_dot is None else _f64(vertices_dot)
        out = np.zeros(max(self.sizes()["nnz"], 1))
        self._check(self._lib.ffh_assemble_jvp(self.handle, quadrature_degree, kind, _ptr(k), float(mass),
                                               _ptr(vd), _ptr(kd), float(mass_dot), _ptr(out)))
        return out[:self.sizes()["nnz"]]

    def assemble_vjp(self, quadrature_degree, values_bar, kappa=1.0, mass=0.0, cell_kappa=None):
        """Cotangent of the H1 or Nedelec matrix values (ffh_assemble_vjp) ->
        (vertices_bar (nv, 3), kappa_bar (scalar, or per cell with
        cell_kappa), mass_bar)."""
        kind, k = self._coefficients(kappa, cell_kappa)
        vb = _f64(values_bar)
        if vb.shape[0] != self.sizes()["nnz"]:
            raise ValueError("values_bar needs one value per stored entry")
        nv = self.nvert
        xb = np.zeros((nv, 3))
        kb = np.zeros(self.ncell if kind == capi.COEFF_CELL else 1)
        mb = C.c_double(0)
        self._check(self._lib.ffh_assemble_vjp(self.handle, quadrature_degree, kind, _ptr(k), float(mass),
                                               _ptr(vb), _ptr(xb), _ptr(kb), C.byref(mb)))
        return xb, (kb if kind == capi.COEFF_CELL else float(kb[0])), mb.value

    def nedelec_pml_values(self, stretch, wave_number, curvilinear=False):
        """Complex PML operator values on the Nedelec pattern
        (ffh_nedelec_pml_values): stretch (T, 3) complex, or (T, 3, 3) with
        curvilinear=True (S[t][i, j] the cell's stretch matrix)."""
        S = np.asarray(stretch, dtype=np.complex128)
        if curvilinear:
            if S.shape != (self.ncell, 3, 3):
                raise ValueError("stretch must be (T, 3, 3)")
            S = np.transpose(S, (0, 2, 1))
        elif S.shape != (self.ncell, 3):
            raise ValueError("stretch must be (T, 3)")
        st = np.ascontiguousarray(np.stack([S.real, S.imag], axis=-1).reshape(-1))
        nnz = self.sizes()["nnz"]
        out = np.zeros(2 * max(nnz, 1))
        self._check(self._lib.ffh_nedelec_pml_values(self.handle, 1 if curvilinear else 0, _ptr(st),
                                                      float(wave_number), _ptr(out)))
        return out[0:2 * nnz:2] + 1j * out[1:2 * nnz:2]

    def nedelec_pml_values_jvp(self, stretch, wave_number, stretch_dot, wave_number_dot,
                               vertices_dot=None, curvilinear=False):
        """Tangent of nedelec_pml_values (ffh_nedelec_pml_values_jvp) along
        stretch_dot (the layout of stretch), wave_number_dot and vertices_dot
        ((nv, 3) or None)."""
        packed = []
        for a in (stretch, stretch_dot):
            S = np.asarray(a, dtype=np.complex128)
            if curvilinear:
                if S.shape != (self.ncell, 3, 3):
                    raise ValueError("stretch must be (T, 3, 3)")
                S = np.transpose(S, (0, 2, 1))
            elif S.shape != (self.ncell, 3):
                raise ValueError("stretch must be (T, 3)")
            packed.append(np.ascontiguousarray(np.stack([S.real, S.imag], axis=-1).reshape(-1)))
        vd = None
        if vertices_dot is not None:
            vd = _f64(np.asarray(vertices_dot, dtype=np.float64).reshape(-1))
            if vd.shape[0] != 3 * self.nvert:
                raise ValueError("vertices_dot must be (nv, 3)")
        nnz = self.sizes()["nnz"]
        out = np.zeros(2 * max(nnz, 1))
        self._check(self._lib.ffh_nedelec_pml_values_jvp(
            self.handle, 1 if curvilinear else 0, _ptr(packed[0]), _ptr(packed[1]), float(wave_number),
            float(wave_number_dot), None if vd is None else _ptr(vd), _ptr(out)))
        return out[0:2 * nnz:2] + 1j * out[1:2 * nnz:2]

    def nedelec_pml_values_vjp(self, stretch, wave_number, values_bar, curvilinear=False):
        """Cotangent of nedelec_pml_values (ffh_nedelec_pml_values_vjp) for
        complex seeds values_bar (nnz) -> (vertices_bar (nv, 3), stretch_bar
        (the stretch layout, complex), wave_number_bar)."""
        S = np.asarray(stretch, dtype=np.complex128)
        shape = (self.ncell, 3, 3) if curvilinear else (self.ncell, 3)
        if S.shape != shape:
            raise ValueError("stretch must be %s" % (shape,))
        if curvilinear:
            S = np.transpose(S, (0, 2, 1))
        st = np.ascontiguousarray(np.stack([S.real, S.imag], axis=-1).reshape(-1))
        nnz = self.sizes()["nnz"]
        w = np.asarray(values_bar, dtype=np.complex128).reshape(-1)
        if w.shape[0] != nnz:
            raise ValueError("values_bar must have nnz entries")
        wb = np.ascontiguousarray(np.stack([w.real, w.imag], axis=-1).reshape(-1))
        vb = np.zeros(3 * self.nvert)
        sb = np.zeros(st.shape[0])
        kb = C.c_double(0.0)
        self._check(self._lib.ffh_nedelec_pml_values_vjp(self.handle, 1 if curvilinear else 0, _ptr(st),
                                                         float(wave_number), _ptr(wb), _ptr(vb), _ptr(sb),
                                                         C.byref(kb)))
        sbar = (sb[0::2] + 1j * sb[1::2]).reshape(S.shape)
        if curvilinear:
            sbar = np.transpose(sbar, (0, 2, 1))
        return vb.reshape(self.nvert, 3), sbar, kb.value

    def load(self, quadrature_degree, kind, params):
        p = _f64(np.atleast_1d(params))
        self._check(self._lib.ffh_load(self.handle, quadrature_degree, kind, _ptr(p), p.size))

    def dirichlet(self, mode=0, values=None):
        vv = None if values is None else _f64(values)
        self._check(self._lib.ffh_dirichlet(self.handle, mode, _ptr(vv)))

    def constrained(self):
        s = self.sizes()
        c = np.zeros(s["n_global"], dtype=np.int32)
        self._check(self._lib.ffh_constrained_download(self.handle, _ptr(c)))
        return c.astype(bool)

    def set_rhs(self, rhs, which=capi.REDUCED):
        r = _f64(rhs)
        self._check(self._lib.ffh_set_rhs(self.handle, which, _ptr(r)))

    def rhs(self, which=capi.FULL):
        s = self.sizes()
        n = s["n_global"] if which == capi.FULL else s["n_free"]
        r = np.zeros(n)
        self._check(self._lib.ffh_rhs_download(self.handle, which, _ptr(r)))
        return r

    def matrix(self, which=capi.FULL) -> "Matrix":
        h = C.c_void_p()
        self._check(self._lib.ffh_problem_matrix(self.handle, which, C.byref(h)))
        return Matrix(h, self.ctx, borrowed=True, owner=self)

    def solution(self):
        """The last solve's solution over all n_global dofs (Dirichlet values included)."""
        x = np.zeros(self.sizes()["n_global"])
        self._check(self._lib.ffh_problem_solution(self.handle, _ptr(x)))
        return x

    def solve(self, opts: SolverOptionsT, download=True):
        s = self.sizes()
        x = np.zeros(s["n_global"]) if download else None
        st = capi.SolverStats()
        self._check(self._lib.ffh_problem_solve(self.handle, C.byref(opts.to_c()), _ptr(x), C.byref(st)))
        method = "cg" if opts.method.strip().lower() == "cg" else "pcg"
        return x, SolverStatsT.from_c(st, method)


class ComplexMatrix:
    """A complex CSC matrix resident in HBM (ffh_zmatrix), in place of the
    reference's retained complex factorisation (fortfem_complex_factor_csc);
    solve() runs Jacobi-COCG, which needs a complex-symmetric matrix (the PML
    curl-curl operators are)."""

    def __init__(self, a: CSC, ctx: Context | None = None, base: int = 1):
        self.ctx = ctx or default_context()
        self._lib = capi.load()
        self.n = int(a.ncol)
        val = np.asarray(a.val, dtype=np.complex128)
        va = np.ascontiguousarray(np.stack([val.real, val.imag], axis=-1).reshape(-1))
        cp, ri = _i32(a.col_ptr), _i32(a.row_idx)
        h = C.c_void_p()
        self.ctx.check(self._lib.ffh_zmatrix_upload(self.ctx.handle, self.n, a.nnz, base, _ptr(cp), _ptr(ri),
                                                    _ptr(va), C.byref(h)))
        self.handle = h

    def solve(self, b, tolerance=1e-10, max_iterations=10000):
        """x with ||b - A x|| <= tolerance ||b|| from x = 0 -> (x, iterations,
        relative residual, converged)."""
        bb = np.asarray(b, dtype=np.complex128).reshape(-1)
        if bb.shape[0] != self.n:
            raise ValueError("b must have n entries")
        bi = np.ascontiguousarray(np.stack([bb.real, bb.imag], axis=-1).reshape(-1))
        x = np.zeros(2 * self.n)
        it, conv = C.c_int(0), C.c_int(0)
        rel = C.c_double(0.0)
        self.ctx.check(self._lib.ffh_zmatrix_solve(self.handle, _ptr(bi), _ptr(x), float(tolerance),
                                                   int(max_iterations), C.byref(it), C.byref(rel), C.byref(conv)))
        return x[0::2] + 1j * x[1::2], it.value, rel.value, bool(conv.value)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.ffh_zmatrix_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """A CSC matrix resident in HBM (ffh_matrix)."""

    def __init__(self, handle, ctx: Context, borrowed=False, owner=None):
        self.handle = handle
        self.ctx = ctx
        self.borrowed = borrowed
        self._owner = owner
        self._lib = capi.load()
        if not borrowed:
            _OWNED.add(self)

    @classmethod
    def upload(cls, a: CSC, ctx: Context | None = None, base: int = 1):
        ctx = ctx or default_context()
        lib = capi.load()
        h = C.c_void_p()
        cp, ri, va = _i32(a.col_ptr), _i32(a.row_idx), _f64(a.val)
        ctx.check(lib.ffh_matrix_upload(ctx.handle, a.ncol, a.nnz, base, _ptr(cp), _ptr(ri), _ptr(va),
                                        C.byref(h)))
        return cls(h, ctx)

    def close(self):
        if getattr(self, "handle", None) and not self.borrowed:
            self._lib.ffh_matrix_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n = C.c_int(0)
        nnz = C.c_int64(0)
        code = self._lib.ffh_matrix_download(self.handle, 1, 0, None, None, None, C.byref(n), C.byref(nnz))
        if code not in (capi.FFH_OK, capi.FFH_ERR_CAPACITY):
            self.ctx.check(code)
        return n.value, nnz.value

    def download(self, base=1) -> CSC:
        n, nnz = self.size()
        cp = np.zeros(n + 1, dtype=np.int32)
        ri = np.zeros(nnz, dtype=np.int32)
        va = np.zeros(nnz)
        nn = C.c_int(0)
        kk = C.c_int64(0)
        self.ctx.check(self._lib.ffh_matrix_download(self.handle, base, nnz, _ptr(cp), _ptr(ri), _ptr(va),
                                                     C.byref(nn), C.byref(kk)))
        return CSC(n, n, cp, ri, va)

    def spmv(self, x):
        n, _ = self.size()
        xx = _f64(x)
        if xx.shape[0] != n:
            raise ValueError("spmv: vector has incompatible size")
        y = np.zeros(n)
        self.ctx.check(self._lib.ffh_spmv(self.handle, _ptr(xx), _ptr(y)))
        return y

    def bench_spmv(self, precision="fp64", reps=50):
        ms = C.c_double(0)
        by = C.c_double(0)
        prec = capi.FP32 if precision == "fp32" else capi.FP64
        self.ctx.check(self._lib.ffh_bench_spmv(self.handle, prec, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def solve(self, b, x, opts: SolverOptionsT):
        """solve_sparse on the resident matrix: x (initial guess) is updated in place."""
        n, _ = self.size()
        bb = _f64(b)
        if bb.shape[0] != n or np.asarray(x).shape[0] != n:
            raise ValueError("solve_sparse: incompatible vector size")
        xx = _f64(x).copy()
        st = capi.SolverStats()
        self.ctx.check(self._lib.ffh_solve_sparse(self.handle, _ptr(bb), _ptr(xx), C.byref(opts.to_c()),
                                                  C.byref(st)))
        x[...] = xx
        method = "cg" if opts.method.strip().lower() == "cg" else "pcg"
        return SolverStatsT.from_c(st, method)

    def spmv_jvp(self, x, values_dot, x_dot):
        """spmv_jvp (fortfem_sparse_matrix.f90:60-91): A x_dot + A_dot x."""
        n, nnz = self.size()
        xx, vd, xd = _f64(x), _f64(values_dot), _f64(x_dot)
        if xx.shape[0] != n:
            raise ValueError("spmv_jvp: vector has incompatible size")
        if xd.shape[0] != n:
            raise ValueError("spmv_jvp: vector_dot has incompatible size")
        if vd.shape[0] != nnz:
            raise ValueError("spmv_jvp: matrix value tangent has incompatible size")
        y = np.zeros(n)
        self.ctx.check(self._lib.ffh_spmv_jvp(self.handle, _ptr(xx), _ptr(vd), _ptr(xd), _ptr(y)))
        return y

    def spmv_vjp(self, x, y_bar):
        """spmv_vjp (fortfem_sparse_matrix.f90:93-126) -> (values_bar, x_bar)."""
        n, nnz = self.size()
        xx, yb = _f64(x), _f64(y_bar)
        if xx.shape[0] != n:
            raise ValueError("spmv_vjp: vector has incompatible size")
        if yb.shape[0] != n:
            raise ValueError("spmv_vjp: product_bar has incompatible size")
        vb = np.zeros(max(nnz, 1))
        xb = np.zeros(n)
        self.ctx.check(self._lib.ffh_spmv_vjp(self.handle, _ptr(xx), _ptr(yb), _ptr(vb), _ptr(xb)))
        return vb[:nnz], xb

    def solve_jvp(self, x, values_dot, b_dot, opts: SolverOptionsT):
        """Tangent of a converged solve (pcg_solve_jvp, advanced_solvers.f90:458-489,
        on the stored pattern): x_dot solves A x_dot = b_dot - A_dot x.
        Returns (x_dot, status, stats); status 0, or 2 when the tangent solve
        did not converge."""
        n, nnz = self.size()
        xx, vd, bd = _f64(x), _f64(values_dot), _f64(b_dot)
        if xx.shape[0] != n or bd.shape[0] != n or vd.shape[0] != nnz:
            raise ValueError("solve_jvp: incompatible sizes")
        xd = np.zeros(n)
        st = capi.SolverStats()
        self.ctx.check(self._lib.ffh_solve_sparse_jvp(self.handle, _ptr(xx), _ptr(vd), _ptr(bd),
                                                      C.byref(opts.to_c()), _ptr(xd), C.byref(st)))
        method = "cg" if opts.method.strip().lower() == "cg" else "pcg"
        stats = SolverStatsT.from_c(st, method)
        return xd, (0 if stats.converged else 2), stats

    def solve_vjp(self, x, x_bar, opts: SolverOptionsT):
        """Cotangent of a converged solve (pcg_solve_vjp, advanced_solvers.f90:491-522):
        lambda solves A^T lambda = x_bar; returns (values_bar, b_bar, status,
        stats) with b_bar = lambda and values_bar = -lambda x^T on the stored
        entries (zeros and status 2 when the adjoint solve did not converge)."""
        n, nnz = self.size()
        xx, xb = _f64(x), _f64(x_bar)
        if xx.shape[0] != n or xb.shape[0] != n:
            raise ValueError("solve_vjp: incompatible sizes")
        vb = np.zeros(max(nnz, 1))
        bb = np.zeros(n)
        st = capi.SolverStats()
        self.ctx.check(self._lib.ffh_solve_sparse_vjp(self.handle, _ptr(xx), _ptr(xb), C.byref(opts.to_c()),
                                                      _ptr(vb), _ptr(bb), C.byref(st)))
        method = "cg" if opts.method.strip().lower() == "cg" else "pcg"
        stats = SolverStatsT.from_c(st, method)
        if not stats.converged:
            return np.zeros(nnz), np.zeros(n), 2, stats
        return vb[:nnz], bb, 0, stats

    def spmv_bytes(self, precision="fp64"):
        """Bytes one SpMV launch moves in the format of the variant in use."""
        by = C.c_double(0)
        prec = capi.FP32 if precision == "fp32" else capi.FP64
        self.ctx.check(self._lib.ffh_matrix_spmv_bytes(self.handle, prec, C.byref(by)))
        return by.value

    def spmv_variant(self, precision="fp64", request=-1):
        """The SpMV variant in use (autotuned on first use); request >= 0 forces one."""
        out = C.c_int(0)
        prec = capi.FP32 if precision == "fp32" else capi.FP64
        self.ctx.check(self._lib.ffh_matrix_spmv_variant(self.handle, prec, request, C.byref(out)))
        return out.value


# ----------------------------------------------------- reference-named calls
def pcg_launches_per_iteration(n: int, precision: str = "fp64") -> int:
    """Launches per Jacobi-PCG iteration on n rows: 3 (direction update,
    SpMV with A p and p.A p, x/r/z update) or 2 (direction formed inside the
    SpMV's gathers, then the update); "fp32" for the mixed solve's
    iterations."""
    lib = capi.load()
    k = lib.ffh_pcg_launches_per_iteration(int(n), capi.FP32 if precision == "fp32" else capi.FP64)
    if k not in (2, 3):
        raise ValueError("pcg_launches_per_iteration: invalid arguments")
    return k


def spmv(matrix: CSC, vector, ctx: Context | None = None):
    """spmv (fortfem_sparse_matrix.f90:46-58): size mismatches raise (error stop)."""
    v = _f64(vector)
    if v.shape[0] != matrix.ncol:
        raise ValueError("spmv: vector has incompatible size")
    return Matrix.upload(matrix, ctx).spmv(v)


def spmv_jvp(matrix: CSC, vector, matrix_values_dot, vector_dot, ctx: Context | None = None):
    """spmv_jvp (fortfem_sparse_matrix.f90:60-91) -> product_dot."""
    if _f64(vector).shape[0] != matrix.ncol:
        raise ValueError("spmv_jvp: vector has incompatible size")
    return Matrix.upload(matrix, ctx).spmv_jvp(vector, matrix_values_dot, vector_dot)


def spmv_vjp(matrix: CSC, vector, product_bar, ctx: Context | None = None):
    """spmv_vjp (fortfem_sparse_matrix.f90:93-126) -> (matrix_values_bar, vector_bar)."""
    if _f64(vector).shape[0] != matrix.ncol:
        raise ValueError("spmv_vjp: vector has incompatible size")
    return Matrix.upload(matrix, ctx).spmv_vjp(vector, product_bar)


def solve_sparse_jvp(a_sparse: CSC, b, x, opts: SolverOptionsT, values_dot, b_dot,
                     ctx: Context | None = None):
    """Sparse pcg_solve_jvp / cg_solve_jvp (advanced_solvers.f90:387-419, 458-489):
    (x_dot, status).  b is the primal right-hand side, kept for the
    converged-state contract as in the reference; it is not read."""
    if a_sparse.ncol != a_sparse.nrow or _f64(b).shape[0] != a_sparse.nrow:
        raise ValueError("solve_sparse_jvp: incompatible sizes")
    xd, status, _ = Matrix.upload(a_sparse, ctx).solve_jvp(x, values_dot, b_dot, opts)
    return xd, status


def solve_sparse_vjp(a_sparse: CSC, b, x, opts: SolverOptionsT, x_bar, ctx: Context | None = None):
    """Sparse pcg_solve_vjp / cg_solve_vjp (advanced_solvers.f90:421-456, 491-522):
    (values_bar, b_bar, status)."""
    if a_sparse.ncol != a_sparse.nrow or _f64(b).shape[0] != a_sparse.nrow:
        raise ValueError("solve_sparse_vjp: incompatible sizes")
    vb, bb, status, _ = Matrix.upload(a_sparse, ctx).solve_vjp(x, x_bar, opts)
    return vb, bb, status


def solve_sparse(a_sparse: CSC, b, x, opts: SolverOptionsT, ctx: Context | None = None):
    """solve_sparse (advanced_solvers.f90:135-195).  x is updated in place and
    the stats are returned.  Non-square or size mismatches raise (error stop)."""
    if a_sparse.ncol != a_sparse.nrow:
        raise ValueError("solve_sparse: only square matrices are supported")
    bb = _f64(b)
    if bb.shape[0] != a_sparse.nrow or np.asarray(x).shape[0] != a_sparse.nrow:
        raise ValueError("solve_sparse: incompatible vector size")
    return Matrix.upload(a_sparse, ctx).solve(bb, x, opts)


def assemble_tetra_lagrange_stiffness_csc(mesh_vertices, tetrahedra, degree, quadrature_degree,
                                          stiffness_coefficient=None, mass_coefficient=None,
                                          cell_coefficient=None, ctx: Context | None = None):
    """assemble_tetra_lagrange_stiffness_csc (:276-334) -> (CSC, status).
    cell_coefficient: optional per-cell stiffness coefficient (new argument)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral H1 sparse assembly failed")
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    if v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status
    if t.min() < 1 or t.max() > v.shape[0]:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_H1, degree)
        p.assemble(quadrature_degree, kappa=1.0 if stiffness_coefficient is None else stiffness_coefficient,
                   mass=0.0 if mass_coefficient is None else mass_coefficient, cell_kappa=cell_coefficient)
        a = p.matrix(capi.FULL).download()
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    return a, FortsparseStatus(0, "")


def assemble_tetra_lagrange_stiffness_csc_jvp(mesh_vertices, tetrahedra, degree, quadrature_degree,
                                              stiffness_coefficient, mass_coefficient, mesh_vertices_dot,
                                              stiffness_coefficient_dot, mass_coefficient_dot,
                                              ctx: Context | None = None):
    """assemble_tetra_lagrange_stiffness_csc_jvp (:336-386) -> (matrix_dot CSC, status).
    mesh_vertices and mesh_vertices_dot are (nv, 3); the result has the
    primal matrix's pattern."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral H1 JVP assembly failed")
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    vd = np.asarray(mesh_vertices_dot)
    if v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status
    if vd.shape != v.shape or t.min() < 1 or t.max() > v.shape[0] or not 1 <= degree <= 4:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_H1, degree)
        vals = p.assemble_jvp(quadrature_degree, kappa=stiffness_coefficient, mass=mass_coefficient,
                              vertices_dot=vd, kappa_dot=stiffness_coefficient_dot, mass_dot=mass_coefficient_dot)
        a = p.matrix(capi.FULL).download()
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    a.val = vals
    return a, FortsparseStatus(0, "")


def assemble_tetra_lagrange_stiffness_csc_vjp(mesh_vertices, tetrahedra, degree, quadrature_degree,
                                              stiffness_coefficient, mass_coefficient, matrix_values_bar,
                                              ctx: Context | None = None):
    """assemble_tetra_lagrange_stiffness_csc_vjp (:388-471) ->
    (mesh_vertices_bar (nv, 3), stiffness_coefficient_bar, mass_coefficient_bar, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral H1 VJP shapes differ")
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    zero = (np.zeros_like(v, dtype=np.float64), 0.0, 0.0)
    if v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return (*zero, status)
    if t.min() < 1 or t.max() > v.shape[0] or not 1 <= degree <= 4:
        return (*zero, status)
    try:
        p = Problem(v, t, ctx)
        _, nnz = p.pattern_build(capi.FAMILY_H1, degree)
        if np.asarray(matrix_values_bar).shape[0] != nnz:
            return (*zero, status)
        xb, kb, mb = p.assemble_vjp(quadrature_degree, matrix_values_bar, kappa=stiffness_coefficient,
                                    mass=mass_coefficient)
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return (*zero, status)
    return xb, kb, mb, FortsparseStatus(0, "")


def nedelec_dof_count(k):
    return k * (k + 2) * (k + 3) // 2


def build_tetra_nedelec_dof_map(order, tetrahedra):
    """build_tetra_nedelec_dof_map (tetra_nedelec_global_dof_map.f90:15-105),
    host routine of the library -> (global_dofs (T, n) 1-based,
    edge_orientations (T, 6), face_permutations (T, 4, 3), ndofs)."""
    t = _i32(np.asarray(tetrahedra).reshape(-1, 4))
    T = t.shape[0]
    n = nedelec_dof_count(order)
    gd = np.zeros((T, n), dtype=np.int32)
    eo = np.zeros((T, 6), dtype=np.int32)
    fp = np.zeros((T, 4, 3), dtype=np.int32)
    nd = C.c_int(0)
    code = capi.load().ffh_nedelec_dof_map(order, T, _ptr(t), 1, _ptr(gd), _ptr(eo), _ptr(fp), C.byref(nd))
    capi.check(code)
    return gd, eo, fp, nd.value


def evaluate_tetra_nedelec_first_kind(order, points):
    """evaluate_tetra_nedelec_first_kind (tetra_nedelec_arbitrary_order.f90:75-106)
    at reference points (npts, 3) -> (values, curls), each (npts, n, 3)."""
    p = _f64(np.atleast_2d(points))
    n = nedelec_dof_count(order)
    vals = np.zeros((p.shape[0], n, 3))
    curls = np.zeros((p.shape[0], n, 3))
    capi.check(capi.load().ffh_nedelec_basis(order, p.shape[0], _ptr(p), _ptr(vals), _ptr(curls)))
    return vals, curls


def assemble_tetra_nedelec_curl_mass_csc(mesh_vertices, tetrahedra, curl_coefficient=None,
                                         mass_coefficient=None, order=None, cell_coefficient=None,
                                         precision="fp64", ctx: Context | None = None):
    """assemble_tetra_nedelec_curl_mass_csc (assembly_tetra_nedelec_3d.f90:
    2042-2112): alpha (curl) and beta (mass) default to 1, order to 1.
    cell_coefficient: optional per-cell curl coefficient (new argument).
    -> (CSC, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral Nedelec sparse assembly failed")
    k = 1 if order is None else int(order)
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    if k < 1 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, k)
        p.assemble(-1, kappa=1.0 if curl_coefficient is None else curl_coefficient,
                   mass=1.0 if mass_coefficient is None else mass_coefficient,
                   cell_kappa=cell_coefficient, precision=precision)
        a = p.matrix(capi.FULL).download()
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    return a, FortsparseStatus(0, "")


def assemble_tetra_nedelec_curl_mass_csc_jvp(mesh_vertices, tetrahedra, order, curl_coefficient,
                                             mass_coefficient, mesh_vertices_dot, curl_coefficient_dot,
                                             mass_coefficient_dot, ctx: Context | None = None):
    """assemble_tetra_nedelec_curl_mass_csc_jvp (assembly_tetra_nedelec_3d.f90:
    2114-2180) -> (matrix_dot CSC on the primal's pattern, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral Nedelec curl-mass JVP assembly failed")
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    vd = np.asarray(mesh_vertices_dot)
    if order < 1 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status
    if vd.shape != v.shape or t.min() < 1 or t.max() > v.shape[0]:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        vals = p.assemble_jvp(-1, kappa=curl_coefficient, mass=mass_coefficient, vertices_dot=vd,
                              kappa_dot=curl_coefficient_dot, mass_dot=mass_coefficient_dot)
        a = p.matrix(capi.FULL).download()
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    a.val = vals
    return a, FortsparseStatus(0, "")


def assemble_tetra_nedelec_curl_mass_csc_vjp(mesh_vertices, tetrahedra, order, curl_coefficient,
                                             mass_coefficient, matrix_values_bar, ctx: Context | None = None):
    """assemble_tetra_nedelec_curl_mass_csc_vjp (:2182-2287) ->
    (mesh_vertices_bar (nv, 3), curl_coefficient_bar, mass_coefficient_bar, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral Nedelec curl-mass VJP shapes differ")
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    zero = (np.zeros_like(v, dtype=np.float64), 0.0, 0.0)
    if order < 1 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return (*zero, status)
    if t.min() < 1 or t.max() > v.shape[0]:
        return (*zero, status)
    try:
        p = Problem(v, t, ctx)
        _, nnz = p.pattern_build(capi.FAMILY_NEDELEC, order)
        if np.asarray(matrix_values_bar).shape[0] != nnz:
            return (*zero, status)
        xb, ab, bb = p.assemble_vjp(-1, matrix_values_bar, kappa=curl_coefficient, mass=mass_coefficient)
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return (*zero, status)
    return xb, ab, bb, FortsparseStatus(0, "")


def _nedelec_pml(mesh_vertices, tetrahedra, order, stretch, wave_number, curvilinear, message, ctx):
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, message)
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    if order < 1 or wave_number <= 0 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        vals = p.nedelec_pml_values(stretch, wave_number, curvilinear)
        a = p.matrix(capi.FULL).download()
    except (FFHError, ValueError) as e:
        if isinstance(e, FFHError) and e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    a.val = vals
    return a, FortsparseStatus(0, "")


def _nedelec_pml_jvp(mesh_vertices, tetrahedra, order, stretch, wave_number, mesh_vertices_dot, stretch_dot,
                     wave_number_dot, curvilinear, message, ctx):
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, message)
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    if order < 1 or wave_number <= 0 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4:
        return None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        vals = p.nedelec_pml_values_jvp(stretch, wave_number, stretch_dot, wave_number_dot,
                                        mesh_vertices_dot, curvilinear)
        a = p.matrix(capi.FULL).download()
    except (FFHError, ValueError) as e:
        if isinstance(e, FFHError) and e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    a.val = vals
    return a, FortsparseStatus(0, "")


def assemble_tetra_nedelec_pml_csc_jvp(mesh_vertices, tetrahedra, order, stretch, wave_number, mesh_vertices_dot,
                                       stretch_dot, wave_number_dot, ctx: Context | None = None):
    """assemble_tetra_nedelec_pml_csc_jvp (assembly_tetra_nedelec_3d.f90:
    1687-1768) -> (matrix_dot: CSC with complex values on the primal pattern,
    status)."""
    return _nedelec_pml_jvp(mesh_vertices, tetrahedra, order, stretch, wave_number, mesh_vertices_dot,
                            stretch_dot, wave_number_dot, False,
                            "Complex tetrahedral Nedelec PML JVP failed", ctx)


def assemble_tetra_nedelec_curvilinear_pml_csc_jvp(mesh_vertices, tetrahedra, order, stretch, wave_number,
                                                   mesh_vertices_dot, stretch_dot, wave_number_dot,
                                                   ctx: Context | None = None):
    """assemble_tetra_nedelec_curvilinear_pml_csc_jvp (:1446-1521); stretch
    and stretch_dot (T, 3, 3) complex."""
    return _nedelec_pml_jvp(mesh_vertices, tetrahedra, order, stretch, wave_number, mesh_vertices_dot,
                            stretch_dot, wave_number_dot, True,
                            "Curvilinear tetrahedral Nedelec PML JVP failed", ctx)


def _nedelec_pml_vjp(mesh_vertices, tetrahedra, order, stretch, wave_number, matrix_values_bar, curvilinear,
                     message, ctx):
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, message)
    v = np.asarray(mesh_vertices)
    t = np.asarray(tetrahedra)
    if order < 1 or wave_number <= 0 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4:
        return None, None, None, status
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        vb, sb, kb = p.nedelec_pml_values_vjp(stretch, wave_number, matrix_values_bar, curvilinear)
    except (FFHError, ValueError) as e:
        if isinstance(e, FFHError) and e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, None, None, status
    return vb, sb, kb, FortsparseStatus(0, "")


def assemble_tetra_nedelec_pml_csc_vjp(mesh_vertices, tetrahedra, order, stretch, wave_number, matrix_values_bar,
                                       ctx: Context | None = None):
    """assemble_tetra_nedelec_pml_csc_vjp (assembly_tetra_nedelec_3d.f90:
    1770-1868): complex seeds on the primal pattern -> (mesh_vertices_bar
    (nv, 3), stretch_bar (T, 3) complex, wave_number_bar, status)."""
    return _nedelec_pml_vjp(mesh_vertices, tetrahedra, order, stretch, wave_number, matrix_values_bar, False,
                            "Complex tetrahedral Nedelec PML VJP failed", ctx)


def assemble_tetra_nedelec_curvilinear_pml_csc_vjp(mesh_vertices, tetrahedra, order, stretch, wave_number,
                                                   matrix_values_bar, ctx: Context | None = None):
    """assemble_tetra_nedelec_curvilinear_pml_csc_vjp (:1523-1611)."""
    return _nedelec_pml_vjp(mesh_vertices, tetrahedra, order, stretch, wave_number, matrix_values_bar, True,
                            "Curvilinear tetrahedral Nedelec PML VJP failed", ctx)


def assemble_tetra_nedelec_pml_csc(mesh_vertices, tetrahedra, order, stretch, wave_number,
                                   ctx: Context | None = None):
    """assemble_tetra_nedelec_pml_csc (assembly_tetra_nedelec_3d.f90:1613-1685):
    stretch (T, 3) complex per cell -> (CSC with complex values (csc_z_t), status)."""
    return _nedelec_pml(mesh_vertices, tetrahedra, order, stretch, wave_number, False,
                        "Complex tetrahedral Nedelec PML assembly failed", ctx)


def assemble_tetra_nedelec_curvilinear_pml_csc(mesh_vertices, tetrahedra, order, stretch, wave_number,
                                               ctx: Context | None = None):
    """assemble_tetra_nedelec_curvilinear_pml_csc (:1372-1444): stretch
    (T, 3, 3) complex per cell -> (CSC with complex values, status)."""
    return _nedelec_pml(mesh_vertices, tetrahedra, order, stretch, wave_number, True,
                        "Curvilinear tetrahedral Nedelec PML assembly failed", ctx)


def assemble_tetra_nedelec_weighted_csc(mesh_vertices, tetrahedra, coefficient, mass_coefficient,
                                        order=None, precision="fp64", ctx: Context | None = None):
    """assemble_tetra_nedelec_weighted_csc (:1870-2000).  coefficient: array
    (T, 3, 3) of symmetric per-cell tensors, (T, nq, 3, 3) tensors at the
    points of nedelec_tensor_points, or a callable K(x, y, z) -> (3, 3),
    evaluated at those points as the reference does (:2497-2501).  Tensors
    constant on every cell use the class-table element; others the per-point
    element (dense curl contraction on MFMA).  -> (CSC, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral weighted Nedelec assembly failed")
    k = 1 if order is None else int(order)
    v = np.asarray(mesh_vertices, dtype=np.float64)
    t = np.asarray(tetrahedra)
    if k < 1 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status
    form, K = _tensor_coefficient(v, t, coefficient, k)
    try:
        p = Problem(v, t, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, k)
        if form == "cell":
            p.assemble(-1, mass=float(mass_coefficient), cell_tensor=K, precision=precision)
        else:
            p.assemble(-1, mass=float(mass_coefficient), point_tensor=K, precision=precision)
        a = p.matrix(capi.FULL).download()
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status
    return a, FortsparseStatus(0, "")


def _physical_points(vertices, tets, quadrature_degree):
    from ._quadrature import tetra_duffy

    x, y, z, _ = tetra_duffy(quadrature_degree)
    v = np.asarray(vertices, dtype=np.float64)
    t = np.asarray(tets) - 1
    v0 = v[t[:, 0]]
    J = np.stack([v[t[:, 1]] - v0, v[t[:, 2]] - v0, v[t[:, 3]] - v0], axis=2)  # (T,3,3) cols
    xi = np.stack([x, y, z], axis=0)  # (3,nq)
    return v0[:, None, :] + np.einsum("tij,jq->tqi", J, xi)


def assemble_tetra_lagrange_scalar_load_samples(mesh_vertices, tetrahedra, degree, quadrature_degree,
                                                source_values, ctx: Context | None = None):
    """assemble_tetra_lagrange_scalar_load_samples (:473-528);
    source_values(nq, T) column-major == array (T, nq) here. -> (rhs, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Sampled tetrahedral H1 scalar load failed")
    try:
        p = Problem(mesh_vertices, tetrahedra, ctx)
        p.pattern_build(capi.FAMILY_H1, degree)
        p.load(quadrature_degree, capi.SOURCE_SAMPLES, _f64(source_values).reshape(-1))
        return p.rhs(capi.FULL), FortsparseStatus(0, "")
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status


def assemble_tetra_lagrange_scalar_load(mesh_vertices, tetrahedra, degree, quadrature_degree, source,
                                        ctx: Context | None = None):
    """assemble_tetra_lagrange_scalar_load (:705-774) with a vectorised
    source(x, y, z); the callback is sampled at the mapped Duffy points."""
    pts = _physical_points(mesh_vertices, tetrahedra, quadrature_degree)
    samples = source(pts[..., 0], pts[..., 1], pts[..., 2])
    return assemble_tetra_lagrange_scalar_load_samples(mesh_vertices, tetrahedra, degree,
                                                       quadrature_degree, samples, ctx)


def nedelec_load_points(order):
    """The reference's Nedelec load quadrature (4-point rule at k = 1, Duffy
    2k + 2 otherwise): array (nq, 4) of (x, y, z, weight)."""
    lib = capi.load()
    nq = C.c_int(0)
    capi.check(lib.ffh_nedelec_load_points(order, C.byref(nq), None))
    pts = np.zeros((nq.value, 4))
    capi.check(lib.ffh_nedelec_load_points(order, C.byref(nq), _ptr(pts)))
    return pts


def assemble_tetra_nedelec_vector_load_samples(mesh_vertices, tetrahedra, source_values, order=1,
                                               ctx: Context | None = None):
    """Nedelec vector load from samples of the source at the mapped load
    points: source_values (T, nq, 3) == f(3, nq, T) column-major.
    -> (vector, status)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral Nedelec vector load failed")
    try:
        p = Problem(mesh_vertices, tetrahedra, ctx)
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        p.load(-1, capi.SOURCE_SAMPLES, _f64(source_values).reshape(-1))
        return p.rhs(capi.FULL), FortsparseStatus(0, "")
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status


def assemble_tetra_nedelec_vector_load(mesh_vertices, tetrahedra, source, order=1,
                                       ctx: Context | None = None):
    """assemble_tetra_nedelec_vector_load (src/assembly/assembly_tetra_nedelec_3d.f90:
    2002-2040) / _order (:71-127) with a vectorised source(x, y, z) returning
    three arrays; the callback is sampled at the mapped load points."""
    samples = _nedelec_source_samples(mesh_vertices, tetrahedra, source, order)
    return assemble_tetra_nedelec_vector_load_samples(mesh_vertices, tetrahedra, samples, order, ctx)


def _nedelec_source_samples(mesh_vertices, tetrahedra, source, order):
    """source(x, y, z) -> (fx, fy, fz) at the mapped load points, (T, nq, 3)."""
    pts = nedelec_load_points(order)
    v = np.asarray(mesh_vertices, dtype=np.float64)
    t = np.asarray(tetrahedra) - 1
    v0 = v[t[:, 0]]
    J = np.stack([v[t[:, 1]] - v0, v[t[:, 2]] - v0, v[t[:, 3]] - v0], axis=2)
    X = v0[:, None, :] + np.einsum("tij,jq->tqi", J, pts[:, :3].T)
    fx, fy, fz = source(X[..., 0], X[..., 1], X[..., 2])
    return np.stack([np.broadcast_to(fx, X.shape[:2]), np.broadcast_to(fy, X.shape[:2]),
                     np.broadcast_to(fz, X.shape[:2])], axis=-1)


def nedelec_tensor_points(mesh_vertices, tetrahedra, order):
    """The points x_q = v1 + J xi_q of the weighted element's rule
    (nedelec_load_points) in every cell: (T, nq, 3)."""
    pts = nedelec_load_points(order)
    v = np.asarray(mesh_vertices, dtype=np.float64)
    t = np.asarray(tetrahedra) - 1
    v0 = v[t[:, 0]]
    J = np.stack([v[t[:, 1]] - v0, v[t[:, 2]] - v0, v[t[:, 3]] - v0], axis=2)
    return v0[:, None, :] + np.einsum("tij,jq->tqi", J, pts[:, :3].T)


def _tensor_coefficient(mesh_vertices, tetrahedra, coefficient, order):
    """The curl tensor of the weighted element as ("cell", (T, 3, 3)) or
    ("point", (T, nq, 3, 3)).  An array is taken as given; a callable
    K(x, y, z) -> (3, 3) is evaluated where the reference evaluates it, at
    every point of the rule mapped into each cell, and the cell form is used
    when it is the same at all points of every cell."""
    t = np.asarray(tetrahedra)
    if not callable(coefficient):
        K = np.asarray(coefficient, dtype=np.float64)
        if K.ndim == 4:
            return "point", K
        return "cell", K.reshape(t.shape[0], 3, 3)
    X = nedelec_tensor_points(mesh_vertices, tetrahedra, order)
    Kp = np.array([[np.asarray(coefficient(*x), dtype=np.float64) for x in cell] for cell in X])
    Kp = Kp.reshape(X.shape[0], X.shape[1], 3, 3)
    if np.array_equal(Kp, np.broadcast_to(Kp[:, :1], Kp.shape)):
        return "cell", np.ascontiguousarray(Kp[:, 0])
    return "point", Kp


def _nedelec_solve(vertices, tetrahedra, order, source, assemble, zero_tangential_boundary, opts, ctx,
                   message):
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, message)
    t = np.asarray(tetrahedra)
    v = np.asarray(vertices)
    if order < 1 or v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 4 or t.shape[0] < 1:
        return None, status, None
    p = Problem(vertices, tetrahedra, ctx)
    try:
        p.pattern_build(capi.FAMILY_NEDELEC, order)
        assemble(p)
        p.load(-1, capi.SOURCE_SAMPLES,
               _f64(_nedelec_source_samples(vertices, tetrahedra, source, order)).reshape(-1))
        if zero_tangential_boundary:
            p.dirichlet(0)
        o = opts or solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-12,
                                   max_iterations=100000, breakdown="relative")
        x, st = p.solve(o)
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status, None
    return x, FortsparseStatus(0, ""), st


def solve_tetra_nedelec_curl_mass(vertices, tetrahedra, order, source, curl_coefficient, mass_coefficient,
                                  zero_tangential_boundary=False, opts: SolverOptionsT | None = None,
                                  ctx: Context | None = None, cell_coefficient=None):
    """solve_tetra_nedelec_curl_mass (fortfem_tetra_nedelec_solver_3d.f90:361-414)
    with the sparse direct solve replaced by Jacobi-PCG (relative 1e-12 unless
    opts says otherwise).  source(x, y, z) -> (fx, fy, fz), vectorised;
    zero_tangential_boundary eliminates the boundary edge and face dofs
    (build_zero_tangential_mask, :416-470) with zero values.
    cell_coefficient: optional per-cell curl coefficient (new argument).
    -> (solution, status, stats)."""
    if mass_coefficient <= 0.0:
        return None, FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral Nedelec curl-mass solve failed"), None
    return _nedelec_solve(
        vertices, tetrahedra, order, source,
        lambda p: p.assemble(-1, kappa=curl_coefficient, mass=mass_coefficient, cell_kappa=cell_coefficient),
        zero_tangential_boundary, opts, ctx, "Tetrahedral Nedelec curl-mass solve failed")


def solve_tetra_nedelec_weighted_curl_mass(vertices, tetrahedra, order, curl_coefficient, source,
                                           mass_coefficient, zero_tangential_boundary=False,
                                           opts: SolverOptionsT | None = None, ctx: Context | None = None):
    """solve_tetra_nedelec_weighted_curl_mass (fortfem_tetra_nedelec_solver_3d.f90:41-95)
    with Jacobi-PCG; curl_coefficient as in assemble_tetra_nedelec_weighted_csc
    (per-cell or per-point tensors, or a callable K(x, y, z) evaluated at the
    rule's points).  -> (solution, status, stats)."""
    if mass_coefficient <= 0.0:
        return None, FortsparseStatus(FORTSPARSE_INVALID_MATRIX,
                                      "Weighted tetrahedral Nedelec curl-mass solve failed"), None
    form, K = _tensor_coefficient(vertices, tetrahedra, curl_coefficient, order)
    kw = {"cell_tensor": K} if form == "cell" else {"point_tensor": K}
    return _nedelec_solve(
        vertices, tetrahedra, order, source,
        lambda p: p.assemble(-1, mass=float(mass_coefficient), **kw),
        zero_tangential_boundary, opts, ctx, "Weighted tetrahedral Nedelec curl-mass solve failed")


def solve_tetra_lagrange_poisson(vertices, tetrahedra, degree, source, boundary_value=None,
                                 opts: SolverOptionsT | None = None, ctx: Context | None = None,
                                 cell_coefficient=None):
    """solve_tetra_lagrange_poisson (fortfem_tetra_lagrange_solver_3d.f90:50-62)
    with the sparse direct solve replaced by Jacobi-PCG.  boundary_value:
    callable g(x, y, z) evaluated at the constrained nodes, or None for g = 0.
    -> (solution, status, stats)."""
    status = FortsparseStatus(FORTSPARSE_INVALID_MATRIX, "Tetrahedral H1 diffusion-reaction solve failed")
    if degree < 1 or degree > 4:
        return None, status, None
    q = 2 * degree + 4
    p = Problem(vertices, tetrahedra, ctx)
    try:
        p.pattern_build(capi.FAMILY_H1, degree)
        p.assemble(q, kappa=1.0, cell_kappa=cell_coefficient)
        pts = _physical_points(vertices, tetrahedra, q)
        p.load(q, capi.SOURCE_SAMPLES, _f64(source(pts[..., 0], pts[..., 1], pts[..., 2])).reshape(-1))
        values = None
        if boundary_value is not None:
            values = _node_values(vertices, tetrahedra, degree, p.global_dofs(base=0), boundary_value)
        p.dirichlet(0, values)
        o = opts or solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                                   max_iterations=100000, breakdown="relative")
        x, st = p.solve(o)
    except FFHError as e:
        if e.code == capi.FFH_ERR_DEVICE:
            raise
        return None, status, None
    return x, FortsparseStatus(0, ""), st


def _node_values(vertices, tets, degree, gd0, g):
    """g at the Lagrange node of every dof (v_1 + J node, as build_dirichlet_data)."""
    from ._quadrature import lagrange_nodes

    nodes = lagrange_nodes(degree)  # (n,3)
    v = np.asarray(vertices, dtype=np.float64)
    t = np.asarray(tets) - 1
    v0 = v[t[:, 0]]
    J = np.stack([v[t[:, 1]] - v0, v[t[:, 2]] - v0, v[t[:, 3]] - v0], axis=2)
    pts = v0[:, None, :] + np.einsum("tij,nj->tni", J, nodes)
    out = np.zeros(int(gd0.max()) + 1)
    out[gd0.reshape(-1)] = g(pts[..., 0], pts[..., 1], pts[..., 2]).reshape(-1)
    return out


__all__ = [
    "CSC", "Context", "FFHError", "FortsparseStatus", "Matrix", "Problem", "SolverOptionsT", "SolverStatsT",
    "assemble_tetra_lagrange_scalar_load", "assemble_tetra_lagrange_scalar_load_samples",
    "assemble_tetra_lagrange_stiffness_csc", "capi", "default_context", "generate_structured_tetra_box_mesh",
    "solve_sparse", "solve_tetra_lagrange_poisson", "solver_options", "spmv", "unit_box_mesh",
    "unit_square_mesh",
]


# ------------------------------------------------------------------ serving
class Server:
    """Continuous-batching solve server (ffh_server_*): requests on prepared
    meshes are admitted by device-memory footprint and advanced together."""

    def __init__(self, budget_bytes=8 << 30, max_systems=64, chunk=16, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._lib = capi.load()
        h = C.c_void_p()
        self.ctx.check(self._lib.ffh_server_create(self.ctx.handle, int(budget_bytes), max_systems, chunk,
                                                   C.byref(h)))
        self.handle = h
        self._n = {}
        _OWNED.add(self)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.ffh_server_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, mesh_key, family, order, vertices, tets):
        v = _f64(vertices)
        t = _i32(tets)
        n = C.c_int(0)
        nnz = C.c_int64(0)
        self.ctx.check(self._lib.ffh_server_prepare(self.handle, mesh_key, family, order, v.shape[0], _ptr(v),
                                                    t.shape[0], _ptr(t), C.byref(n), C.byref(nnz)))
        self._n[(mesh_key, family, order)] = (n.value, t.shape[0])
        return n.value, nnz.value

    def submit(self, mesh_key, family, order, cell_coeff, rhs, mass=0.0, tol=1e-10, max_iterations=100000):
        """rhs: (n_free,) or (nrhs, n_free) array.  Returns a ticket."""
        b = _f64(np.atleast_2d(rhs))
        c = _f64(cell_coeff)
        tk = C.c_int64(0)
        self.ctx.check(self._lib.ffh_server_submit(self.handle, mesh_key, family, order, _ptr(c), float(mass),
                                                   b.shape[0], _ptr(b), float(tol), int(max_iterations),
                                                   C.byref(tk)))
        return tk.value, b.shape[0], b.shape[1]

    def step(self):
        a, q, f = C.c_int(0), C.c_int(0), C.c_int(0)
        self.ctx.check(self._lib.ffh_server_step(self.handle, C.byref(a), C.byref(q), C.byref(f)))
        return a.value, q.value, f.value

    def wait(self, ticket, nrhs, n, max_steps=1000000):
        """Steps the server until the request has finished (ffh_server_wait);
        the results as poll(), or None if max_steps were not enough."""
        x = np.zeros((nrhs, n))
        it = np.zeros(nrhs, dtype=np.int32)
        cv = np.zeros(nrhs, dtype=np.int32)
        rr = np.zeros(nrhs)
        code = self._lib.ffh_server_wait(self.handle, ticket, max_steps, _ptr(x), _ptr(it), _ptr(cv), _ptr(rr))
        if code == 1:
            return None
        self.ctx.check(code)
        return x, it, cv.astype(bool), rr

    def poll(self, ticket, nrhs, n):
        """None while pending, else (solutions (nrhs, n), iterations, converged, rel_residual)."""
        x = np.zeros((nrhs, n))
        it = np.zeros(nrhs, dtype=np.int32)
        cv = np.zeros(nrhs, dtype=np.int32)
        rr = np.zeros(nrhs)
        code = self._lib.ffh_server_poll(self.handle, ticket, _ptr(x), _ptr(it), _ptr(cv), _ptr(rr))
        if code == 1:
            return None
        self.ctx.check(code)
        return x, it, cv.astype(bool), rr

    def use_graph(self, enable=True):
        """Replay each step's chunk from a HIP graph (ffh_server_use_graph)."""
        self.ctx.check(self._lib.ffh_server_use_graph(self.handle, 1 if enable else 0))

    def set_tiling(self, mode=-1):
        """Batched SpMV per system type: -1 autotune (default), 0 plain SELL-64,
        1 LDS column tiles wherever they fit, 2 / 4 / 8 row-split with that
        many lanes per row wherever R allows it (ffh_server_set_tiling)."""
        self.ctx.check(self._lib.ffh_server_set_tiling(self.handle, mode))

    def tiling_stats(self):
        """(system types chosen so far, how many run the tiled product)."""
        n, t = C.c_int(0), C.c_int(0)
        self.ctx.check(self._lib.ffh_server_tiling_stats(self.handle, C.byref(n), C.byref(t)))
        return n.value, t.value

    def product_forms(self):
        """(system types chosen so far, how many run the tiled product, how
        many a row-split product) (ffh_server_product_forms)."""
        n, t, l = C.c_int(0), C.c_int(0), C.c_int(0)
        self.ctx.check(self._lib.ffh_server_product_forms(self.handle, C.byref(n), C.byref(t),
                                                          C.byref(l)))
        return n.value, t.value, l.value

    def set_precision(self, policy=-1):
        """0 FP32 iterations (switch to FP64 on stagnation), 1 FP64, 2 FP64 for
        Nedelec; returns the number of requests switched so far."""
        sw = C.c_int64(0)
        self.ctx.check(self._lib.ffh_server_set_precision(self.handle, policy, C.byref(sw)))
        return sw.value

    def progress(self, ticket, nrhs):
        st = np.zeros(nrhs, dtype=np.int32)
        it = np.zeros(nrhs, dtype=np.int32)
        rr = np.zeros(nrhs)
        self.ctx.check(self._lib.ffh_server_progress(self.handle, ticket, _ptr(st), _ptr(it), _ptr(rr)))
        return st, it, rr

    def stats(self):
        u, i = C.c_int64(0), C.c_int64(0)
        self.ctx.check(self._lib.ffh_server_stats(self.handle, C.byref(u), C.byref(i)))
        return {"used_bytes": u.value, "iterations_run": i.value}

    def run_until_done(self, tickets, max_steps=1000000):
        """Steps until every (ticket, nrhs, n) has finished; returns {ticket: result}."""
        out = {}
        pending = list(tickets)
        for _ in range(max_steps):
            if not pending:
                break
            self.step()
            rest = []
            for (tk, nrhs, n) in pending:
                r = self.poll(tk, nrhs, n)
                if r is None:
                    rest.append((tk, nrhs, n))
                else:
                    out[tk] = r
            pending = rest
        if pending:
            raise RuntimeError("server did not finish the requests")
        return out

