"""Nedelec solver drivers on the GPU.

solve_tetra_nedelec_curl_mass (src/solvers/fortfem_tetra_nedelec_solver_3d.f90:
361-414) and solve_tetra_nedelec_weighted_curl_mass (:41-95) with the sparse
direct solve replaced by Jacobi-PCG (relative 1e-12).

Checks: the reference's own tests (a constant source on the two-tet mesh is
reproduced by every order 1..6 to 3e-10 (k <= 4) / 2e-8 (k = 5, 6),
test/test_public_tetra_nedelec_arbitrary_order_solve_slow.f90:26-46; the
order-5 manufactured curl-curl field on the six-tet cube with zero tangential
trace is within 0.2 of the exact field at (0.31, 0.27, 0.42),
test/test_curl_curl_solution_field.f90:26-71), and the solution against a
dense direct solve of the oracle's system (the reference solves directly) to
1e-8 relative, with and without the boundary elimination, scalar and
per-cell tensor coefficients.
"""
import numpy as np
import pytest

import oracle as O
from test_gpu_nedelec_parity import _tangential_mask

pytestmark = pytest.mark.gpu

TWO_TETS = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], dtype=np.int32)
TWO_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
CONST = np.array([1.25, -0.75, 0.5])


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def cube_mesh():
    """test_curl_curl_solution_field.f90:111-143: six tets around the
    diagonal, odd-oriented cells flipped."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]],
                 dtype=np.float64)
    t = np.array([[1, 2, 4, 8], [1, 2, 6, 8], [1, 3, 4, 8], [1, 3, 7, 8], [1, 5, 6, 8], [1, 5, 7, 8]],
                 dtype=np.int32)
    for c in range(6):
        x = v[t[c] - 1]
        if np.linalg.det(np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)) < 0:
            t[c, 2], t[c, 3] = t[c, 3], t[c, 2]
    return v, t


def barycentric_cell(v, t, point):
    for c in range(t.shape[0]):
        x = v[t[c] - 1]
        J = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
        xi = np.linalg.solve(J, point - x[0])
        if xi.min() >= -1e-12 and xi.sum() <= 1 + 1e-12:
            return c, xi
    raise AssertionError("point outside the mesh")


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_constant_field_every_order(ff, k):
    x, st, stats = ff.solve_tetra_nedelec_curl_mass(TWO_V, TWO_TETS, k, lambda a, b, c: tuple(CONST), 1.0, 1.0)
    assert st.code == 0 and stats.converged
    tol = 3.0e-10 if k <= 4 else 2.0e-8
    err = 0.0
    for cell in range(2):
        for xi in ([0.25, 0.25, 0.25], [0.1, 0.2, 0.3], [0.6, 0.1, 0.1], [0.05, 0.05, 0.85]):
            err = max(err, np.max(np.abs(O.nedelec_field(k, TWO_V, TWO_TETS, x, cell, np.array(xi)) - CONST)))
    assert err < tol


def test_manufactured_curl_curl_order5(ff):
    v, t = cube_mesh()

    def source(x, y, z):
        return (0.0, 0.0, (2 * np.pi ** 2 + 1.0) * np.sin(np.pi * x) * np.sin(np.pi * y))

    x, st, stats = ff.solve_tetra_nedelec_curl_mass(v, t, 5, source, 1.0, 1.0, zero_tangential_boundary=True)
    assert st.code == 0 and stats.converged
    point = np.array([0.31, 0.27, 0.42])
    cell, xi = barycentric_cell(v, t, point)
    value = O.nedelec_field(5, v, t, x, cell, xi)
    exact = np.array([0.0, 0.0, np.sin(np.pi * point[0]) * np.sin(np.pi * point[1])])
    assert np.max(np.abs(value - exact)) < 0.2


def _dense_reference(k, v, t, a, b, cons):
    A = a.to_dense()
    if cons is None:
        return np.linalg.solve(A, b)
    free = ~cons
    x = np.zeros(A.shape[0])
    x[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    return x


def _source(x, y, z):
    return (np.sin(x + 2 * y) + 0.3, z * z - x, np.cos(3 * z) * y)


@pytest.mark.parametrize("k,boundary", [(1, False), (2, False), (2, True), (3, True)])
def test_curl_mass_matches_dense_solve(ff, k, boundary):
    v, t = O.unit_box_mesh(2)
    v = v + 0.04 * np.random.default_rng(k).uniform(-1, 1, v.shape)
    a = O.nedelec_csc(k, v, t, alpha=1.7, beta=0.6)
    b = O.nedelec_vector_load(k, v, t, _source)
    cons = _tangential_mask(k, v, t) if boundary else None
    xr = _dense_reference(k, v, t, a, b, cons)
    x, st, stats = ff.solve_tetra_nedelec_curl_mass(v, t, k, _source, 1.7, 0.6, zero_tangential_boundary=boundary)
    assert st.code == 0 and stats.converged
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


@pytest.mark.parametrize("boundary", [False, True])
def test_weighted_curl_mass_matches_dense_solve(ff, boundary):
    k = 2
    v, t = O.unit_box_mesh(2)
    rng = np.random.default_rng(7)
    L = rng.uniform(-0.3, 0.3, (t.shape[0], 3, 3))
    K = np.einsum("tij,tkj->tik", L, L) + np.eye(3) * 1.2  # SPD per cell
    a = O.nedelec_csc(k, v, t, beta=0.8, cell_tensor=K)
    b = O.nedelec_vector_load(k, v, t, _source)
    cons = _tangential_mask(k, v, t) if boundary else None
    xr = _dense_reference(k, v, t, a, b, cons)
    x, st, stats = ff.solve_tetra_nedelec_weighted_curl_mass(v, t, k, K, _source, 0.8,
                                                            zero_tangential_boundary=boundary)
    assert st.code == 0 and stats.converged
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


def test_rejects_invalid_input(ff):
    x, st, _ = ff.solve_tetra_nedelec_curl_mass(TWO_V, TWO_TETS, 1, lambda a, b, c: tuple(CONST), 1.0, 0.0)
    assert x is None and st.code != 0
    x, st, _ = ff.solve_tetra_nedelec_curl_mass(TWO_V, TWO_TETS, 0, lambda a, b, c: tuple(CONST), 1.0, 1.0)
    assert x is None and st.code != 0
