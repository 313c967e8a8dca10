"""Pins the 2D P1 part of the oracle (configuration C1): the element against
the textbook gradient formula, the load against the exact integral of a
constant, and the unit-source Poisson solution against its Fourier series.

Reference: src/fortfem_forms_simple.f90:1678-1761 (element, load),
src/solvers/fortfem_solvers_laplacian.f90:88-97 (symmetric elimination),
src/fortfem_api_mesh.f90:128-139 (unit_square_mesh).
"""
import numpy as np

import oracle as O


def series_centre(terms=2001):
    """u(1/2, 1/2) of -lap u = 1 on the unit square, u = 0 on the boundary:
    sum over odd m, n of 16 sin(m pi/2) sin(n pi/2) / (pi^4 m n (m^2 + n^2))."""
    m = np.arange(1, terms, 2, dtype=float)
    M, N = np.meshgrid(m, m, indexing="ij")
    s = np.sin(M * np.pi / 2) * np.sin(N * np.pi / 2)
    return float(np.sum(16 * s / (np.pi ** 4 * M * N * (M * M + N * N))))


def single_sum_centre(terms=41):
    """The same value from the single-series form
    u = x(1-x)/2 - sum_{m odd} 4 sin(m pi x) cosh(m pi (y - 1/2)) / (pi^3 m^3 cosh(m pi / 2))."""
    m = np.arange(1, terms, 2, dtype=float)
    return float(0.125 - np.sum(4 * np.sin(m * np.pi / 2) / (np.pi ** 3 * m ** 3 * np.cosh(m * np.pi / 2))))


def test_series_value_matches_known_constant():
    # the centre value 0.07367135328 also appears in the reference's
    # magnetic-box test (test/test_paper_magnetic_box_3d.f90:8)
    assert abs(single_sum_centre() - 0.07367135328) < 1e-11
    assert abs(series_centre() - single_sum_centre()) < 1e-10


def test_single_triangle_element_and_load():
    v = np.array([[0.2, 0.1], [1.3, 0.4], [0.5, 1.7]])
    t = np.array([[1, 2, 3]], dtype=np.int32)
    A = O.p1_2d_dense(v, t, mass=0.0, stiffness=2.5)
    x, y = v[:, 0], v[:, 1]
    area = 0.5 * abs((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]])
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]])
    K = 2.5 * (np.outer(b, b) + np.outer(c, c)) / (4 * area)
    assert np.max(np.abs(A - K)) < 1e-14 * np.max(np.abs(K))
    M = O.p1_2d_dense(v, t, mass=1.0, stiffness=0.0)
    assert np.max(np.abs(M - area / 12 * (np.ones((3, 3)) + np.eye(3)))) < 1e-15
    assert np.allclose(O.p1_2d_load(v, t, 3.0), 3.0 * area / 3)


def _centre_value(n):
    a, b, (v, t, bnd) = O.p1_2d_poisson_system(n)
    x = np.linalg.solve(a.to_dense(), b)
    c = np.argmin(np.sum((v - 0.5) ** 2, axis=1))
    return x[c], a, b, bnd


def test_unit_source_converges_at_second_order():
    exact = single_sum_centre()
    e33 = abs(_centre_value(33)[0] - exact)
    e65, a, b, bnd = _centre_value(65)
    e65 = abs(e65 - exact)
    assert e65 < 3e-5
    assert 3.5 < e33 / e65 < 4.5
    # symmetric elimination leaves identity rows and columns at boundary dofs
    d = a.to_dense()
    i = np.flatnonzero(bnd)[0]
    assert d[i, i] == 1.0 and np.count_nonzero(d[i]) == 1 and np.count_nonzero(d[:, i]) == 1
    assert np.all(b[bnd] == 0.0)


def test_pcg_matches_direct_solve_c1():
    a, b, _ = O.p1_2d_poisson_system(65)
    x, st = O.solve_sparse(a, b, method="pcg", preconditioner="jacobi", tolerance=1e-10,
                           max_iterations=5000, breakdown="relative")
    xd = np.linalg.solve(a.to_dense(), b)
    assert st.converged
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
