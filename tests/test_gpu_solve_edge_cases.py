"""solve_sparse edge cases against the oracle's restatement of the reference
loop (advanced_solvers.f90:135-195, 831-925, 1063-1086), on the device path:

* a 1 x 1 system (one row, one nonzero);
* a zero right-hand side: the relative tolerance is then zero, the initial
  residual meets it and the solve stops at once with x unchanged;
* a diagonal system: Jacobi-PCG is exact after one iteration;
* an initial guess that already solves the system;
* a system whose rows range from 1 to several hundred nonzeros.

Iteration counts are compared exactly with the oracle's; solutions to
1e-12 (relative) where the oracle stops in the same iteration.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def _solve_both(ff, A, b, x0=None, prec="fp64"):
    a = O.sparse_from_dense(A)
    n = A.shape[0]
    m = ff.Matrix.upload(ff.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, a.val))
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=float)
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=1000, breakdown="relative", precision=prec)
    st = m.solve(b, x, opts)
    xo, sto = O.solve_sparse(a, b, x0=None if x0 is None else np.array(x0, dtype=float), method="pcg",
                             preconditioner="jacobi", tolerance=1e-10, max_iterations=1000,
                             breakdown="relative")
    return x, st, xo, sto


def _rel(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


def test_one_by_one(ff):
    x, st, xo, sto = _solve_both(ff, np.array([[4.0]]), np.array([2.0]))
    assert st.converged and sto.converged and st.iterations == sto.iterations
    assert abs(x[0] - 0.5) <= 1e-15


def test_zero_rhs_stops_at_once(ff):
    A = np.diag(np.arange(1.0, 6.0)) + 0.1 * (np.eye(5, k=1) + np.eye(5, k=-1))
    x, st, xo, sto = _solve_both(ff, A, np.zeros(5))
    assert st.iterations == sto.iterations == 0
    assert st.converged == sto.converged
    assert np.all(x == 0.0)


def test_diagonal_system_one_iteration(ff):
    d = np.linspace(1.0, 9.0, 40)
    b = np.cos(np.arange(40.0))
    for prec in ("fp64", "mixed"):
        x, st, xo, sto = _solve_both(ff, np.diag(d), b, prec=prec)
        assert st.converged
        if prec == "fp64":
            assert st.iterations == sto.iterations
            assert _rel(x, b / d) <= 1e-14
        else:
            assert _rel(x, b / d) <= 1e-8


def test_exact_initial_guess(ff):
    rng = np.random.default_rng(2)
    M = rng.normal(size=(12, 12))
    A = M @ M.T + 12 * np.eye(12)
    xs = rng.normal(size=12)
    b = A @ xs
    x, st, xo, sto = _solve_both(ff, A, b, x0=xs)
    assert st.converged and st.iterations == sto.iterations
    assert _rel(x, xs) <= 1e-13


def test_ragged_rows(ff):
    """rows of 1 to 301 nonzeros: an arrow with a dense first row/column,
    a tridiagonal band in the middle and isolated diagonal rows at the end"""
    n = 400
    A = np.zeros((n, n))
    A[0, 1:301] = A[1:301, 0] = 0.01
    for i in range(1, 300):
        A[i, i + 1] = A[i + 1, i] = -0.3
    A[np.arange(n), np.arange(n)] = 2.0 + np.arange(n) / n
    b = np.sin(np.arange(n, dtype=float))
    x, st, xo, sto = _solve_both(ff, A, b)
    assert st.converged and sto.converged
    assert abs(st.iterations - sto.iterations) <= 1
    assert _rel(x, np.linalg.solve(A, b)) <= 1e-9
