"""The threaded forms of the oracle's assembly and solve stages (used only to
time the CPU baseline on all host cores) against the sequential oracle: the
triplet stage and csc_from_triplet give the same CSC bit for bit at any
thread count; the threaded Jacobi-PCG agrees to rounding (its dot products
are summed per thread)."""
import numpy as np
import pytest

import oracle as O
from fortfem_amd import workloads


@pytest.mark.parametrize("d,N", [(1, 6), (2, 4)])
def test_parallel_assembly_equals_sequential(d, N):
    v, t = O.unit_box_mesh(N)
    kappa = workloads.lognormal_kappa(t.shape[0], 4)
    q = 2 * d + 4
    a, g, G = O.lagrange_stiffness_csc(v, t, d, q, cell_kappa=kappa)
    for nt in (1, 3, 8):
        b = O.par_lagrange_stiffness_csc(v, t, d, q, g, G, nt, cell_kappa=kappa)
        assert np.array_equal(a.col_ptr, b.col_ptr) and np.array_equal(a.row_idx, b.row_idx)
        assert np.array_equal(a.val, b.val)


def test_parallel_pcg_agrees_with_sequential():
    v, t = O.unit_box_mesh(8)
    kappa = workloads.lognormal_kappa(t.shape[0], 4)
    a, g, G = O.lagrange_stiffness_csc(v, t, 1, 6, cell_kappa=kappa)
    cons, _ = O.lagrange_dirichlet(v, t, 1, g, G)
    red, free, _ = O.constrained_reduction(a, cons, np.zeros(G), np.zeros(G))
    b = workloads.uniform_rhs(red.n_rows, 4)
    xs, ss = O.solve_sparse(red, b, tolerance=1e-10, max_iterations=5000, breakdown="relative")
    for nt in (1, 4):
        xp, sp = O.par_solve_sparse(red, b, nt, tolerance=1e-10, max_iterations=5000, breakdown="relative")
        assert sp.converged and abs(sp.iterations - ss.iterations) <= 2
        assert np.linalg.norm(xp - xs) <= 1e-8 * np.linalg.norm(xs)
