"""Every SpMV variant (CSR stream, vector CSR with 4..32 lanes per row, SELL-64,
SELL-64 with LDS column tiles, several grid sizes) against the oracle's matrix-vector product and PCG solve.

Tolerances: FP64 products within 1e-13 of the largest entry of the oracle's
product (only the summation order differs); FP32 products within 1e-5;
solutions within 1e-9 of the oracle's (the north-star solution bar is 1e-12
for the assembled values; PCG stops at a 1e-10 relative residual).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
NV = 13


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def _laplace_p2(N):
    v, t = O.unit_box_mesh(N)
    a, g, G = O.lagrange_stiffness_csc(v, t, 2, 6)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    b = np.ones(G)
    red, free, rb = O.constrained_reduction(a, cons, np.zeros(G), b)
    return red, rb


def _dense_row_matrix(n=3000):
    """SPD arrow matrix: row/column 0 is full (longer than one LDS stage)."""
    rng = np.random.default_rng(7)
    A = np.zeros((n, n))
    A[0, 1:] = A[1:, 0] = rng.uniform(-1, 1, n - 1) / n
    A[np.arange(n), np.arange(n)] = 2.0 + rng.uniform(0, 1, n)
    for i in range(1, n - 1):
        A[i, i + 1] = A[i + 1, i] = -0.5
    return O.sparse_from_dense(A), A


def test_variant_count_and_bounds(ff):
    red, rb = _laplace_p2(2)
    m = ff.Matrix.upload(ff.CSC(red.n_rows, red.n_cols, red.col_ptr, red.row_idx, red.val))
    assert 0 <= m.spmv_variant("fp64") < NV
    assert 0 <= m.spmv_variant("fp32") < NV
    with pytest.raises(ff.FFHError):
        m.spmv_variant("fp64", NV)


@pytest.mark.parametrize("case", ["p2_laplace", "dense_row"])
def test_every_variant_matches_oracle_product(ff, case):
    if case == "p2_laplace":
        a, _ = _laplace_p2(10)
    else:
        a, _ = _dense_row_matrix()
    m = ff.Matrix.upload(ff.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, a.val))
    x = np.random.default_rng(3).uniform(-1, 1, a.n_rows)
    y_ref = a.matvec(x)
    scale = np.max(np.abs(y_ref))
    for vi in range(NV):
        assert m.spmv_variant("fp64", vi) == vi
        y = m.spmv(x)
        assert np.max(np.abs(y - y_ref)) <= 1e-13 * scale, vi
        ms, by = m.bench_spmv("fp64", reps=3)
        assert ms > 0 and by > 0
        m.spmv_variant("fp32", vi)
        ms, by = m.bench_spmv("fp32", reps=3)
        assert ms > 0 and by > 0


@pytest.mark.parametrize("case", ["p2_laplace", "dense_row"])
def test_every_variant_solves_like_oracle(ff, case):
    if case == "p2_laplace":
        a, b = _laplace_p2(10)
    else:
        a, A = _dense_row_matrix()
        b = A @ np.linspace(1, 2, a.n_rows)
    xo, sto = O.solve_sparse(a, b, method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    m = ff.Matrix.upload(ff.CSC(a.n_rows, a.n_cols, a.col_ptr, a.row_idx, a.val))
    for vi in range(NV):
        m.spmv_variant("fp64", vi)
        m.spmv_variant("fp32", vi)
        for prec in ("fp64", "mixed"):
            for graph in (False, True):
                x = np.zeros(a.n_rows)
                opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                                         max_iterations=5000, breakdown="relative", precision=prec,
                                         use_graph=graph)
                st = m.solve(b, x, opts)
                assert st.converged, (vi, prec, graph)
                if prec == "fp64":
                    assert abs(st.iterations - sto.iterations) <= 1, (vi, graph)
                assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo), (vi, prec, graph)
