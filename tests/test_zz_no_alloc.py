"""The allocation-free solve (ffh_solve_sparse_device_noalloc) against
ffh_solve_sparse, through the C-ABI.

The caller's device arrays are made with the HIP runtime the library itself
links (libamdhip64.so.7, the same process-wide instance), so the test
allocates and the entry point under test does not: the free device memory is
the same before and after each call.  Two calls in a row, the second on a
smaller system, each bit-exact against ffh_solve_sparse on the same matrix
with the same SpMV variant (solve_sparse, src/solvers/advanced_solvers.f90:
135-195, Jacobi-PCG :831-925).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        from fortfem_amd import _capi

        _capi.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def download(self, p: C.c_void_p, a: np.ndarray) -> np.ndarray:
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0
        return a

    def free_bytes(self) -> int:
        f, t = C.c_size_t(0), C.c_size_t(0)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def release(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _system(N: int, degree: int, seed: int):
    """The reduced Jacobi-PCG system of P<degree> Poisson on the N^3 x 6 box
    with lognormal per-cell kappa (exactly symmetric: every element writes
    its upper triangle to both halves), 0-based, and a U[-1, 1] rhs."""
    import fortfem_amd as ff
    from fortfem_amd import workloads

    v, t = ff.unit_box_mesh(N)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, degree)
    p.assemble(2 * degree + 4, cell_kappa=workloads.lognormal_kappa(t.shape[0], seed))
    p.dirichlet(0)
    a = p.matrix(ff.capi.REDUCED).download(base=0)
    b = workloads.uniform_rhs(a.nrow, seed)
    return a, b


def _opts(ff, **kw):
    base = dict(method="pcg", preconditioner="jacobi", tolerance=1e-10, max_iterations=20000,
                breakdown="relative", check_interval=16)
    base.update(kw)
    return ff.solver_options(**base)


def _noalloc(ff, hip, a, b, opts):
    lib = ff.capi.load()
    ctx = ff.default_context()
    cp, ri, va = hip.upload(a.col_ptr), hip.upload(a.row_idx), hip.upload(a.val)
    db, dx = hip.upload(np.ascontiguousarray(b)), hip.upload(np.zeros(a.nrow))
    st = ff.capi.SolverStats()
    before = hip.free_bytes()
    code = lib.ffh_solve_sparse_device_noalloc(ctx.handle, a.nrow, a.nnz, cp, ri, va, db, dx,
                                               C.byref(opts.to_c()), C.byref(st))
    after = hip.free_bytes()
    x = hip.download(dx, np.zeros(a.nrow))
    return code, x, st, before, after


def test_noalloc_solve_matches_solve_sparse_twice():
    import fortfem_amd as ff

    lib = ff.capi.load()
    hip = _Hip()
    try:
        # a larger, then a smaller system: the array keeps the first call's
        # contents, which must not leak into the second
        for N, d, seed in ((14, 2, 4), (6, 1, 2)):
            print(f"noalloc case N={N} P{d}", flush=True)
            a, b = _system(N, d, seed)
            opts = _opts(ff)
            code, x, st, before, after = _noalloc(ff, hip, a, b, opts)
            assert code == 0, lib.ffh_last_error(ff.default_context().handle)
            assert before == after, (before, after)  # nothing allocated by the call
            vi = lib.ffh_noalloc_spmv_variant(a.nrow, a.nnz)
            m = ff.Matrix.upload(a, base=0)
            chosen = C.c_int(-1)
            assert lib.ffh_matrix_spmv_variant(m.handle, ff.capi.FP64, vi, C.byref(chosen)) == 0
            assert chosen.value == vi
            x_ref = np.zeros(a.nrow)
            st_ref = m.solve(b, x_ref, opts)
            assert st.converged == 1 and st_ref.converged
            assert st.iterations == st_ref.iterations
            assert np.array_equal(x, x_ref)  # same kernels, same launch shapes: bit for bit
            assert st.final_residual == st_ref.final_residual
            m.close()
            hip.release()
    finally:
        hip.release()


def test_noalloc_solve_refusals():
    """Sizes beyond the array, a nonsymmetric matrix (checked on the device
    by a bounded kernel) and mixed precision are refused with -1."""
    import fortfem_amd as ff

    lib = ff.capi.load()
    ctx = ff.default_context()
    hip = _Hip()
    try:
        a, b = _system(3, 1, 2)
        cp, ri, va = hip.upload(a.col_ptr), hip.upload(a.row_idx), hip.upload(a.val)
        db, dx = hip.upload(b), hip.upload(np.zeros(a.nrow))
        st = ff.capi.SolverStats()
        o = _opts(ff).to_c()
        big = ff.capi.NOALLOC_MAX_ROWS + 1
        assert lib.ffh_solve_sparse_device_noalloc(ctx.handle, big, a.nnz, cp, ri, va, db, dx,
                                                   C.byref(o), C.byref(st)) == -1
        om = _opts(ff, precision="mixed").to_c()
        assert lib.ffh_solve_sparse_device_noalloc(ctx.handle, a.nrow, a.nnz, cp, ri, va, db, dx,
                                                   C.byref(om), C.byref(st)) == -1
        # break the symmetry of one off-diagonal entry
        v2 = a.val.copy()
        j = 0
        k = int(a.col_ptr[j])
        while int(a.row_idx[k]) == j:
            k += 1
        v2[k] *= 1.5
        va2 = hip.upload(v2)
        assert lib.ffh_solve_sparse_device_noalloc(ctx.handle, a.nrow, a.nnz, cp, ri, va2, db, dx,
                                                   C.byref(o), C.byref(st)) == -1
        # and the intact matrix still solves after the refusals
        assert lib.ffh_solve_sparse_device_noalloc(ctx.handle, a.nrow, a.nnz, cp, ri, va, db, dx,
                                                   C.byref(o), C.byref(st)) == 0
        assert st.converged == 1
    finally:
        hip.release()
