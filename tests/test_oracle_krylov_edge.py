"""The oracle's Jacobi-PCG on degenerate inputs, against the reference loop
(advanced_solvers.f90:831-925, 1063-1086) followed by hand: with b = 0 and
x0 = 0 the residual, z and p are zero, the first p.Ap is 0 and the loop
exits on the breakdown test; finalize_solver_status then marks the solve
converged (0 <= tol = 0) with 0 iterations and x unchanged -- in the
reference's absolute breakdown mode and in the relative one alike."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("mode", ["absolute", "relative"])
@pytest.mark.parametrize("method", ["pcg", "cg"])
def test_zero_rhs_converges_at_once(mode, method):
    A = np.diag(np.arange(1.0, 6.0)) + 0.1 * (np.eye(5, k=1) + np.eye(5, k=-1))
    x, st = O.solve_sparse(O.sparse_from_dense(A), np.zeros(5), method=method, preconditioner="jacobi",
                           tolerance=1e-10, max_iterations=50, breakdown=mode)
    assert st.converged and st.iterations == 0
    assert np.all(x == 0.0)


def test_one_by_one_is_one_iteration():
    x, st = O.solve_sparse(O.sparse_from_dense(np.array([[4.0]])), np.array([2.0]), method="pcg",
                           preconditioner="jacobi", tolerance=1e-10, max_iterations=50,
                           breakdown="relative")
    assert st.converged and st.iterations == 1 and x[0] == 0.5
