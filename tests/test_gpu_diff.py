"""Differentiable variants on the GPU against the oracle.

spmv_jvp / spmv_vjp (src/solvers/fortfem_sparse_matrix.f90:60-126) and the
sparse pcg/cg solve tangents and adjoints
(src/solvers/advanced_solvers.f90:387-522) through the C-ABI
(ffh_spmv_jvp, ffh_spmv_vjp, ffh_solve_sparse_jvp, ffh_solve_sparse_vjp).

Bars: the products round differently from the oracle's column loop (the
JVP sums each row in ascending column order with fused multiply-adds), so
they agree to 1e-13 relative and repeat bit for bit from call to call; the
cotangent values are one product each and agree exactly.  The solve
tangents are converged Krylov solves, compared with the oracle's at
tolerance 1e-12 to within 1e-8 relative (FP64 and mixed), and checked with
the reference tests' own central differences and adjoint identities on
their 3x3 fixtures.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# test/test_sparse_matrix_ad.f90:19-28
SPMV_A = np.array([4.0, -1.0, 0.5, 0.7, 3.0, -0.4, -0.2, 0.8, 2.5]).reshape(3, 3, order="F")
SPMV_AD = np.array([0.12, -0.03, 0.05, -0.07, 0.09, 0.02, 0.04, -0.06, 0.08]).reshape(3, 3, order="F")
SPMV_X = np.array([0.6, -0.8, 0.35])
SPMV_XD = np.array([-0.11, 0.07, 0.09])
SPMV_YB = np.array([0.31, -0.27, 0.44])
# test/test_pcg_solver_ad.f90:18-29
PCG_A = np.array([5.0, 1.0, 0.0, 1.0, 4.0, 0.3, 0.0, 0.3, 2.5]).reshape(3, 3, order="F")
PCG_AD = np.array([0.09, -0.02, 0.01, -0.02, 0.05, 0.03, 0.01, 0.03, -0.06]).reshape(3, 3, order="F")
PCG_B = np.array([0.8, -0.25, 0.55])
PCG_BD = np.array([-0.06, 0.11, 0.04])
PCG_XB = np.array([0.22, -0.35, 0.41])
# test/test_cg_solver_ad.f90:18-29
CG_A = np.array([4.0, 1.0, 0.0, 1.0, 3.0, 0.2, 0.0, 0.2, 2.0]).reshape(3, 3, order="F")
CG_AD = np.array([0.11, -0.03, 0.02, -0.03, 0.07, 0.01, 0.02, 0.01, -0.04]).reshape(3, 3, order="F")
CG_B = np.array([1.0, -0.4, 0.7])
CG_BD = np.array([-0.08, 0.12, 0.05])
CG_XB = np.array([0.27, -0.31, 0.44])


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def full_csc(dense):
    n = dense.shape[0]
    cp = (np.arange(n + 1) * n + 1).astype(np.int32)
    ri = np.tile(np.arange(1, n + 1), n).astype(np.int32)
    return O.CSC(n, n, cp, ri, np.asarray(dense, dtype=np.float64).reshape(-1, order="F").copy())


def pattern_values(a, dense):
    cols = np.repeat(np.arange(a.n_cols), np.diff(a.col_ptr))
    return dense[a.row_idx - 1, cols]


def with_values(a, vals):
    return O.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, np.asarray(vals, dtype=np.float64))


def upload(ff, a):
    return ff.Matrix.upload(ff.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, a.val))


def random_nonsymmetric(n, density, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < density) + np.diag(rng.uniform(2, 3, n))
    return O.sparse_from_dense(d)


def p2_system(N=6, seed=0):
    v, t = O.unit_box_mesh(N)
    kappa = np.exp(np.random.default_rng(seed).normal(size=t.shape[0]))
    a, g, G = O.lagrange_stiffness_csc(v, t, 2, 8, cell_kappa=kappa)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    red, _, _ = O.constrained_reduction(a, cons, np.zeros(G), np.zeros(G))
    return red


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_spmv_ad_reference_fixture(ff):
    a = O.sparse_from_dense(SPMV_A)
    m = upload(ff, a)
    vd = pattern_values(a, SPMV_AD)
    yd = m.spmv_jvp(SPMV_X, vd, SPMV_XD)
    assert rel(yd, O.spmv_jvp(a, SPMV_X, vd, SPMV_XD)) <= 1e-15
    step = 2.0e-7
    yp = upload(ff, O.sparse_from_dense(SPMV_A + step * SPMV_AD)).spmv(SPMV_X + step * SPMV_XD)
    ym = upload(ff, O.sparse_from_dense(SPMV_A - step * SPMV_AD)).spmv(SPMV_X - step * SPMV_XD)
    fd = (yp - ym) / (2 * step)
    assert np.max(np.abs(yd - fd)) / max(1.0, np.max(np.abs(yd))) < 2.0e-8
    vb, xb = m.spmv_vjp(SPMV_X, SPMV_YB)
    vbo, xbo = O.spmv_vjp(a, SPMV_X, SPMV_YB)
    assert np.array_equal(vb, vbo)
    assert rel(xb, xbo) <= 1e-15
    lhs = SPMV_YB @ yd
    rhs = xb @ SPMV_XD + vb @ vd
    assert abs(lhs - rhs) < 2.0e-10 * max(1.0, abs(lhs), abs(rhs))


@pytest.mark.parametrize("kind", ["nonsymmetric", "p2_poisson"])
def test_spmv_ad_matches_oracle(ff, kind):
    a = random_nonsymmetric(700, 0.01, 1) if kind == "nonsymmetric" else p2_system()
    rng = np.random.default_rng(2)
    x, xd, yb = rng.uniform(-1, 1, (3, a.n_cols))
    vd = rng.uniform(-1, 1, a.nnz)
    m = upload(ff, a)
    yd = m.spmv_jvp(x, vd, xd)
    assert rel(yd, O.spmv_jvp(a, x, vd, xd)) <= 1e-13
    assert np.array_equal(yd, m.spmv_jvp(x, vd, xd))  # deterministic: no atomics
    vb, xb = m.spmv_vjp(x, yb)
    vbo, xbo = O.spmv_vjp(a, x, yb)
    assert np.array_equal(vb, vbo)
    assert rel(xb, xbo) <= 1e-13
    # the primal product is untouched by the tangent calls
    assert rel(m.spmv(x), O.csc_matvec(a, x)) <= 1e-13


def test_spmv_ad_rejects_wrong_sizes(ff):
    a = O.sparse_from_dense(SPMV_A)
    m = upload(ff, a)
    with pytest.raises(ValueError):
        m.spmv_jvp(SPMV_X, np.zeros(a.nnz + 1), SPMV_XD)
    with pytest.raises(ValueError):
        m.spmv_vjp(SPMV_X, SPMV_YB[:2])


@pytest.mark.parametrize("which", ["pcg", "cg"])
def test_solve_ad_reference_fixture(ff, which):
    A, Ad, b, bd, xb, precond = ((PCG_A, PCG_AD, PCG_B, PCG_BD, PCG_XB, "jacobi") if which == "pcg"
                                 else (CG_A, CG_AD, CG_B, CG_BD, CG_XB, "none"))
    opts = ff.solver_options(method=which, preconditioner=precond, tolerance=1e-12,
                             tolerance_type="absolute", max_iterations=40)
    a = full_csc(A)
    m = upload(ff, a)
    x = np.zeros(3)
    assert m.solve(b, x, opts).converged
    vd = pattern_values(a, Ad)
    xd, status, _ = m.solve_jvp(x, vd, bd, opts)
    step = 2.0e-7
    xp = np.zeros(3)
    upload(ff, full_csc(A + step * Ad)).solve(b + step * bd, xp, opts)
    xm = np.zeros(3)
    upload(ff, full_csc(A - step * Ad)).solve(b - step * bd, xm, opts)
    fd = (xp - xm) / (2 * step)
    assert status == 0
    assert np.max(np.abs(xd - fd)) / max(1.0, np.max(np.abs(xd))) < 2.0e-7
    vb, bb, status, _ = m.solve_vjp(x, xb, opts)
    assert status == 0
    lhs = xb @ xd
    rhs = vb @ vd + bb @ bd
    assert abs(lhs - rhs) < 3.0e-8 * max(1.0, abs(lhs), abs(rhs))
    kw = dict(method=which, preconditioner=precond, tolerance=1e-12, tolerance_type="absolute",
              max_iterations=40)
    xo, _ = O.solve_sparse(a, b, np.zeros(3), **kw)
    assert rel(xd, O.solve_sparse_jvp(a, xo, vd, bd, **kw)[0]) <= 1e-9
    vbo, bbo, _, _ = O.solve_sparse_vjp(a, xo, xb, **kw)
    assert rel(bb, bbo) <= 1e-9 and rel(vb, vbo) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_ad_p2_poisson_matches_oracle(ff, precision):
    a = p2_system()
    rng = np.random.default_rng(5)
    b, bd, xb = rng.uniform(-1, 1, (3, a.n_cols))
    vd = rng.uniform(-0.05, 0.05, a.nnz) * np.abs(a.val)
    kw = dict(method="pcg", preconditioner="jacobi", tolerance=1e-12, max_iterations=20000,
              breakdown="relative")
    xo, sto = O.solve_sparse(a, b, **kw)
    assert sto.converged
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-12, max_iterations=20000,
                             precision=precision, breakdown="relative")
    m = upload(ff, a)
    xd, status, st = m.solve_jvp(xo, vd, bd, opts)
    xdo, _, _ = O.solve_sparse_jvp(a, xo, vd, bd, **kw)
    assert status == 0 and st.converged
    assert np.linalg.norm(xd - xdo) <= 1e-8 * np.linalg.norm(xdo)
    vb, bb, status, st = m.solve_vjp(xo, xb, opts)
    vbo, bbo, _, _ = O.solve_sparse_vjp(a, xo, xb, **kw)
    assert status == 0 and st.converged
    assert np.linalg.norm(bb - bbo) <= 1e-8 * np.linalg.norm(bbo)
    assert np.linalg.norm(vb - vbo) <= 1e-8 * np.linalg.norm(vbo)
    lhs = xb @ xd
    assert abs(lhs - (vb @ vd + bb @ bd)) <= 1e-8 * max(1.0, abs(lhs))


def test_solve_ad_status_and_errors(ff):
    a = p2_system()
    m = upload(ff, a)
    ones = np.ones(a.n_cols)
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-14, max_iterations=1)
    _, status, st = m.solve_jvp(ones, np.ones(a.nnz), ones, opts)
    assert status == 2 and not st.converged
    vb, bb, status, _ = m.solve_vjp(ones, ones, opts)
    assert status == 2 and not vb.any() and not bb.any()
    ns = upload(ff, random_nonsymmetric(50, 0.1, 3))
    good = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10, max_iterations=100)
    with pytest.raises(Exception):
        ns.solve_jvp(np.ones(50), np.ones(ns.size()[1]), np.ones(50), good)
    with pytest.raises(Exception):
        ns.solve_vjp(np.ones(50), np.ones(50), good)
