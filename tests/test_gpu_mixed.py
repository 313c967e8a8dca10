"""Mixed-precision solve (FFH_MIXED of solve_sparse): FP32 SELL-64 iterations
with FP64 residual refreshes, checked against the oracle's FP64 Jacobi-PCG
(advanced_solvers.f90:831-925) from the same initial guess.

Bars: converged with the FP64 true residual ||b - A x|| <= tol ||b - A x0||
(relative) or <= tol (absolute), and the solution within 1e-8 of the
oracle's (the mixed-precision bar).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def _p2_system(N=6, seed=0):
    v, t = O.unit_box_mesh(N)
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(size=t.shape[0]))
    a, g, G = O.lagrange_stiffness_csc(v, t, 2, 8, cell_kappa=kappa)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    red, free, _ = O.constrained_reduction(a, cons, np.zeros(G), np.zeros(G))
    return red


def _upload(ff, a):
    return ff.Matrix.upload(ff.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, a.val))


def test_mixed_from_nonzero_initial_guess(ff):
    a = _p2_system()
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, a.n_rows)
    x0 = rng.uniform(-1, 1, a.n_rows)
    xo, sto = O.solve_sparse(a, b, x0=x0.copy(), method="pcg", preconditioner="jacobi", tolerance=1e-12,
                             max_iterations=20000, breakdown="relative")
    m = _upload(ff, a)
    x = x0.copy()
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10, max_iterations=20000,
                             precision="mixed")
    st = m.solve(b, x, opts)
    r0 = np.linalg.norm(b - a.to_dense() @ x0)
    assert st.converged
    assert abs(st.initial_residual - r0) <= 1e-12 * r0
    assert st.true_residual <= 1e-10 * r0
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_mixed_absolute_tolerance(ff):
    a = _p2_system(seed=2)
    b = np.random.default_rng(3).uniform(-1, 1, a.n_rows)
    tol = 1e-9 * np.linalg.norm(b)
    m = _upload(ff, a)
    x = np.zeros(a.n_rows)
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=tol, max_iterations=20000,
                             precision="mixed", tolerance_type="absolute")
    st = m.solve(b, x, opts)
    assert st.converged and st.true_residual <= tol
    assert np.linalg.norm(b - a.to_dense() @ x) <= 1.01 * tol


def test_mixed_moves_to_fp64_when_fp32_stalls(ff):
    """Unpreconditioned CG on a matrix with condition number ~1e9: FP32
    iterations cannot resolve a 1e-10 residual, the FP64 refreshes stop
    improving and the solve continues in FP64 to the requested tolerance."""
    n = 120
    d = np.logspace(0, 9, n)
    A = np.diag(d)
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = 0.25 * min(d[i], d[i + 1])
    a = O.sparse_from_dense(A)
    xe = np.random.default_rng(4).uniform(-1, 1, n)
    b = A @ xe
    m = _upload(ff, a)
    x = np.zeros(n)
    opts = ff.solver_options(method="cg", tolerance=1e-10, max_iterations=100000, precision="mixed")
    st = m.solve(b, x, opts)
    assert st.converged and st.true_residual <= 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
