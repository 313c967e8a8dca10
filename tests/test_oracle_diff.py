"""Oracle restatement of the differentiable variants, pinned on the CPU.

spmv_jvp / spmv_vjp (src/solvers/fortfem_sparse_matrix.f90:60-126) and the
sparse forms of pcg_solve_jvp/vjp and cg_solve_jvp/vjp
(src/solvers/advanced_solvers.f90:387-522).  The 3x3 fixtures, steps and
bars are those of the reference's own tests (test/test_sparse_matrix_ad.f90,
test/test_pcg_solver_ad.f90, test/test_cg_solver_ad.f90); the random cases
add independent routes: the JVP is bilinear, so a central difference with
step 1 is exact, and the solve tangent has the closed form
A^-1 (b_dot - A_dot x) through a dense LU.
"""
import numpy as np
import pytest

import oracle as O

# test/test_sparse_matrix_ad.f90:19-28 (column-major reshape)
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


def full_csc(dense):
    """All n^2 entries stored (explicit zeros kept), so a dense tangent maps
    one-to-one onto the value array."""
    n = dense.shape[0]
    cp = (np.arange(n + 1) * n + 1).astype(np.int32)
    ri = np.tile(np.arange(1, n + 1), n).astype(np.int32)
    return O.CSC(n, n, cp, ri, np.asarray(dense, dtype=np.float64).reshape(-1, order="F").copy())


def pattern_values(a, dense):
    """The entries of a dense matrix at a CSC pattern (the test's loop at
    test_sparse_matrix_ad.f90:31-37)."""
    cols = np.repeat(np.arange(a.n_cols), np.diff(a.col_ptr))
    return dense[a.row_idx - 1, cols]


def with_values(a, vals):
    return O.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, np.asarray(vals, dtype=np.float64))


def random_sparse(n, density, seed, symmetric=False, spd=False):
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < density)
    if symmetric or spd:
        d = np.triu(d) + np.triu(d, 1).T
    if spd:
        d = d + np.diag(np.abs(d).sum(axis=1) + 1.0)
    else:
        d = d + np.diag(rng.uniform(2, 3, n))
    return O.sparse_from_dense(d)


def test_spmv_ad_reference_fixture():
    a = O.sparse_from_dense(SPMV_A)
    vd = pattern_values(a, SPMV_AD)
    yd = O.spmv_jvp(a, SPMV_X, vd, SPMV_XD)
    step = 2.0e-7
    yp = O.csc_matvec(O.sparse_from_dense(SPMV_A + step * SPMV_AD), SPMV_X + step * SPMV_XD)
    ym = O.csc_matvec(O.sparse_from_dense(SPMV_A - step * SPMV_AD), SPMV_X - step * SPMV_XD)
    fd = (yp - ym) / (2 * step)
    assert np.max(np.abs(yd - fd)) / max(1.0, np.max(np.abs(yd))) < 2.0e-8
    vb, xb = O.spmv_vjp(a, SPMV_X, SPMV_YB)
    lhs = SPMV_YB @ yd
    rhs = xb @ SPMV_XD + vb @ vd
    assert abs(lhs - rhs) < 2.0e-10 * max(1.0, abs(lhs), abs(rhs))


@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_ad_nonsymmetric_exact_difference(seed):
    """Bilinearity: (A+Ad)(x+xd) - (A-Ad)(x-xd) = 2 (A xd + Ad x) exactly in
    exact arithmetic; a transposed operand, a dropped term or a sign error
    fails it on a nonsymmetric matrix."""
    a = random_sparse(60, 0.08, seed)
    rng = np.random.default_rng(10 + seed)
    x, xd, yb = rng.uniform(-1, 1, (3, a.n_cols))
    vd = rng.uniform(-1, 1, a.nnz)
    yd = O.spmv_jvp(a, x, vd, xd)
    fd = (O.csc_matvec(with_values(a, a.val + vd), x + xd) - O.csc_matvec(with_values(a, a.val - vd), x - xd)) / 2
    assert np.max(np.abs(yd - fd)) <= 1e-14 * max(1.0, np.max(np.abs(fd)))
    vb, xb = O.spmv_vjp(a, x, yb)
    # x_bar = A^T y_bar through the dense transpose; values_bar by the
    # identity one entry at a time (unit tangents)
    assert np.allclose(xb, a.to_dense().T @ yb, rtol=0, atol=1e-13)
    for k in rng.choice(a.nnz, 8, replace=False):
        e = np.zeros(a.nnz)
        e[k] = 1.0
        assert abs(yb @ O.spmv_jvp(a, x, e, np.zeros(a.n_cols)) - vb[k]) <= 1e-15 * max(1.0, abs(vb[k]))
    lhs = yb @ yd
    rhs = xb @ xd + vb @ vd
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_spmv_ad_rejects_wrong_sizes():
    a = O.sparse_from_dense(SPMV_A)
    with pytest.raises(ValueError):
        O.spmv_jvp(a, SPMV_X, np.zeros(a.nnz + 1), SPMV_XD)
    with pytest.raises(ValueError):
        O.spmv_vjp(a, SPMV_X[:2], SPMV_YB)


def _solve_ad_fixture(A, Ad, b, bd, xb, method, precond):
    a = full_csc(A)
    opts = dict(method=method, preconditioner=precond, tolerance=1e-12, tolerance_type="absolute",
                max_iterations=40)
    x, st = O.solve_sparse(a, b, np.zeros(3), **opts)
    assert st.converged
    vd = pattern_values(a, Ad)
    xd, _, status = O.solve_sparse_jvp(a, x, vd, bd, **opts)
    step = 2.0e-7
    xp, _ = O.solve_sparse(full_csc(A + step * Ad), b + step * bd, np.zeros(3), **opts)
    xm, _ = O.solve_sparse(full_csc(A - step * Ad), b - step * bd, np.zeros(3), **opts)
    fd = (xp - xm) / (2 * step)
    assert status == 0
    assert np.max(np.abs(xd - fd)) / max(1.0, np.max(np.abs(xd))) < 2.0e-7
    # closed form through LU
    assert np.allclose(xd, np.linalg.solve(A, bd - Ad @ x), rtol=0, atol=1e-11)
    vb, bb, _, status = O.solve_sparse_vjp(a, x, xb, **opts)
    assert status == 0
    # the reference's dense A_bar = -lambda x^T, read on the full pattern
    assert np.allclose(vb, pattern_values(a, -np.outer(bb, x)), rtol=0, atol=0)
    lhs = xb @ xd
    rhs = vb @ vd + bb @ bd
    assert abs(lhs - rhs) < 3.0e-8 * max(1.0, abs(lhs), abs(rhs))


def test_pcg_ad_reference_fixture():
    _solve_ad_fixture(PCG_A, PCG_AD, PCG_B, PCG_BD, PCG_XB, "pcg", "jacobi")


def test_cg_ad_reference_fixture():
    _solve_ad_fixture(CG_A, CG_AD, CG_B, CG_BD, CG_XB, "cg", "none")


def test_solve_ad_sparse_spd_closed_form():
    """Sparse SPD system: x_dot = A^-1 (b_dot - A_dot x) with A_dot on the
    stored entries, and the adjoint identity <x_bar, x_dot> =
    <values_bar, values_dot> + <b_bar, b_dot>."""
    a = random_sparse(80, 0.06, 3, spd=True)
    rng = np.random.default_rng(4)
    b, bd, xb = rng.uniform(-1, 1, (3, a.n_cols))
    vd = rng.uniform(-0.1, 0.1, a.nnz)
    opts = dict(method="pcg", preconditioner="jacobi", tolerance=1e-13, max_iterations=2000,
                breakdown="relative")
    x, st = O.solve_sparse(a, b, **opts)
    assert st.converged
    xd, _, status = O.solve_sparse_jvp(a, x, vd, bd, **opts)
    A = a.to_dense()
    Ad = with_values(a, vd).to_dense()
    assert status == 0
    ref = np.linalg.solve(A, bd - Ad @ x)
    assert np.linalg.norm(xd - ref) <= 1e-11 * np.linalg.norm(ref)
    vb, bb, _, status = O.solve_sparse_vjp(a, x, xb, **opts)
    assert status == 0
    assert np.linalg.norm(bb - np.linalg.solve(A.T, xb)) <= 1e-11 * np.linalg.norm(bb)
    lhs = xb @ xd
    rhs = vb @ vd + bb @ bd
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))


def test_solve_ad_reports_unconverged_status():
    a = random_sparse(80, 0.06, 5, spd=True)
    b = np.ones(a.n_cols)
    x, _ = O.solve_sparse(a, b, tolerance=1e-13, max_iterations=2000, breakdown="relative")
    opts = dict(tolerance=1e-14, max_iterations=1)
    _, _, status = O.solve_sparse_jvp(a, x, np.ones(a.nnz), b, **opts)
    assert status == 2
    vb, bb, _, status = O.solve_sparse_vjp(a, x, b, **opts)
    assert status == 2 and not vb.any() and not bb.any()


def test_identity_catches_a_planted_sign_error():
    """A values_bar with the wrong sign (a plausible slip of the -lambda x^T
    rule) fails the adjoint identity the tests above rely on."""
    a = random_sparse(40, 0.1, 6, spd=True)
    rng = np.random.default_rng(7)
    b, bd, xb = rng.uniform(-1, 1, (3, a.n_cols))
    vd = rng.uniform(-0.1, 0.1, a.nnz)
    x, _ = O.solve_sparse(a, b, tolerance=1e-13, max_iterations=2000, breakdown="relative")
    xd, _, _ = O.solve_sparse_jvp(a, x, vd, bd, tolerance=1e-13, max_iterations=2000, breakdown="relative")
    vb, bb, _, _ = O.solve_sparse_vjp(a, x, xb, tolerance=1e-13, max_iterations=2000, breakdown="relative")
    lhs = xb @ xd
    assert abs(lhs - (vb @ vd + bb @ bd)) <= 1e-10 * max(1.0, abs(lhs))
    assert abs(lhs - (-vb @ vd + bb @ bd)) > 1e-6 * max(1.0, abs(lhs))
