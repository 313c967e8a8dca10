"""GPU parity of the 2D P1 path (configuration C1) against the oracle:
unit_square_mesh(65), f = 1, symmetric elimination of the boundary
(fortfem_solvers_laplacian.f90:88-97), sparse_from_dense's drop of
|a| <= 1e-14 (fortfem_sparse_matrix.f90:16-44), Jacobi-PCG to 1e-10.

Tolerances: reduced pattern bit-exact, values and right-hand side within
1e-12 of the largest entry, solution within 1e-9 (PCG stops at 1e-10).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


@pytest.mark.parametrize("n,f,g,mass", [(65, 1.0, 0.0, 0.0), (17, 2.0, 0.5, 0.3)])
def test_c1_system_and_solution(ff, n, f, g, mass):
    v, t, bnd = ff.unit_square_mesh(n)
    p = ff.Problem(v, t)
    G, _ = p.pattern_build(ff.capi.FAMILY_P1_2D, 1)
    assert G == n * n
    from colouring import assert_valid_colouring

    assert_valid_colouring(p)
    p.assemble(0, kappa=1.0, mass=mass)
    p.load(0, ff.capi.SOURCE_CONSTANT, [f])
    p.dirichlet(1, values=np.full(G, g))
    red = p.matrix(ff.capi.REDUCED).download()
    rhs = p.rhs(ff.capi.REDUCED)
    # oracle: dense assembly, symmetric elimination, sparse_from_dense
    vo, to, bo = O.unit_square_mesh(n)
    a = O.p1_2d_dense(vo, to, mass, 1.0)
    b = O.p1_2d_load(vo, to, f)
    a, b = O.p1_2d_symmetric_dirichlet(a, b, bo, g)
    ref = O.sparse_from_dense(a)
    assert np.array_equal(red.col_ptr, ref.col_ptr) and np.array_equal(red.row_idx, ref.row_idx)
    assert np.max(np.abs(red.val - ref.val)) <= 1e-12 * np.max(np.abs(ref.val))
    assert np.max(np.abs(rhs - b)) <= 1e-12 * np.max(np.abs(b))
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    x, st = p.solve(opts)
    xo, sto = O.solve_sparse(ref, b, method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    assert st.converged and abs(st.iterations - sto.iterations) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    if n == 65:
        # unit source: the centre value converges to the series solution
        c = np.argmin(np.sum((v - 0.5) ** 2, axis=1))
        assert abs(x[c] - 0.07367135328) < 3e-5
