"""Pins the CPU oracle of the tetrahedral H1 path to the reference's own
known-answer tests and to closed forms that a plausible mistake would break.

Reference tests restated: test/test_structured_tetra_box_mesh.f90:19-51,
test/test_tetra_duffy_quadrature.f90:13-33,
test/test_tetra_lagrange_arbitrary_order.f90:19-70,
test/test_tetra_lagrange_global_dof_map.f90:15-37,
test/test_tetra_lagrange_arbitrary_order_assembly.f90:28-66,
test/test_tetra_lagrange_poisson_p_convergence.f90:26-40.
"""
from fractions import Fraction
from itertools import product
from math import factorial

import numpy as np
import pytest

import oracle as O


# ----------------------------------------------------------------- meshes --
def test_box_mesh_reference_kat():
    bounds = [[-1.0, 2.0, 4.0], [1.0, 5.0, 10.0]]
    counts = [2, 1, 3]
    v, t = O.box_mesh(bounds, counts)
    assert v.shape[0] == np.prod(np.array(counts) + 1)
    assert t.shape[0] == 6 * np.prod(counts)
    assert np.all(np.abs(v.min(0) - bounds[0]) < 1e-14) and np.all(np.abs(v.max(0) - bounds[1]) < 1e-14)
    v0 = v[t[:, 0] - 1]
    J = np.stack([v[t[:, k] - 1] - v0 for k in (1, 2, 3)], axis=2)
    det = np.linalg.det(J)
    assert np.all(det > 0)
    assert abs(det.sum() / 6 - np.prod(np.array(bounds[1]) - bounds[0])) < 1e-12


def test_box_mesh_closed_form_ids():
    """vertex id = 1 + i + (n1+1)(j + (n2+1)k); every tet contains brick corners 1 and 8."""
    counts = [3, 2, 2]
    v, t = O.box_mesh([[0, 0, 0], [1, 1, 1]], counts)
    n1, n2 = counts[0] + 1, counts[1] + 1
    for vid in (1, 7, 20, v.shape[0]):
        i = (vid - 1) % n1
        j = ((vid - 1) // n1) % n2
        k = (vid - 1) // (n1 * n2)
        assert np.allclose(v[vid - 1], [i / counts[0], j / counts[1], k / counts[2]])
    first = t.reshape(-1, 6, 4)
    assert np.all(first[:, :, 0] == first[:, :1, 0])  # corner 1 shared by the six


def test_unit_square_boundary_matches_edge_count():
    """find_boundary (mesh_2d.f90:553-603): endpoints of single-triangle edges."""
    v, t, bnd = O.unit_square_mesh(7)
    from collections import Counter

    c = Counter()
    for tri in t:
        for k in range(3):
            a, b = sorted((tri[k], tri[(k + 1) % 3]))
            c[(a, b)] += 1
    brute = np.zeros(v.shape[0], dtype=bool)
    for (a, b), n in c.items():
        if n == 1:
            brute[a - 1] = brute[b - 1] = True
    assert np.array_equal(brute, bnd)
    assert bnd.sum() == 4 * 6


# ------------------------------------------------------------- quadrature --
def test_gauss_legendre_against_golub_welsch():
    for n in range(1, 12):
        x, w = O.gauss_legendre_ab(n, 0.0, 1.0)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - 0.5 * (xr + 1))) < 2e-15
        assert np.max(np.abs(w - 0.5 * wr)) < 2e-15


def test_duffy_reference_kat():
    for p in range(8):
        x, y, z, w = O.tetra_duffy(p)
        assert abs(w.sum() - 1 / 6) < 2e-14
        for a in range(p + 1):
            for b in range(p + 1 - a):
                for c in range(p + 1 - a - b):
                    exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                    assert abs(np.sum(w * x ** a * y ** b * z ** c) - exact) < 4e-14
    with pytest.raises(ValueError):
        O.tetra_duffy(-1)


def test_triangle_duffy_exactness():
    for p in range(8):
        x, y, w = O.triangle_duffy(p)
        for a in range(p + 1):
            for b in range(p + 1 - a):
                exact = factorial(a) * factorial(b) / factorial(a + b + 2)
                assert abs(np.sum(w * x ** a * y ** b) - exact) < 4e-15


# ------------------------------------------------------------------ basis --
@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4])
def test_lagrange_basis_reference_kat(degree):
    assert O.lagrange_dof_count(degree) == [1, 4, 10, 20, 35][degree]
    bary, nodes = O.lagrange_init(degree)
    for k, node in enumerate(nodes):
        st, vals, _ = O.lagrange_eval(degree, node)
        vals[k] -= 1.0
        assert st == 0 and np.max(np.abs(vals)) < 3e-13
    pt = np.array([0.17, 0.21, 0.13])
    st, vals, grads = O.lagrange_eval(degree, pt)
    assert abs(vals.sum() - 1) < 3e-13 and np.max(np.abs(grads.sum(0))) < 2e-12
    if degree >= 1:
        f = 2 * nodes[:, 0] - 3 * nodes[:, 1] + 0.5 * nodes[:, 2] + 1.25
        assert abs(vals @ f - (2 * pt[0] - 3 * pt[1] + 0.5 * pt[2] + 1.25)) < 3e-13
    for c in range(3):
        e = np.zeros(3)
        e[c] = 1e-6
        _, vp, _ = O.lagrange_eval(degree, pt + e)
        _, vm, _ = O.lagrange_eval(degree, pt - e)
        assert np.max(np.abs((vp - vm) / 2e-6 - grads[:, c])) < 2e-7


def test_lagrange_local_order():
    """P1 local order is (v4, v3, v2, v1); P2 = v4, e34, v3, e24, e23, v2, e14, e13, e12, v1."""
    bary, nodes = O.lagrange_init(1)
    assert np.array_equal(bary, [[0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]])
    bary, _ = O.lagrange_init(2)
    assert bary[1].tolist() == [0, 0, 1, 1] and bary[-1].tolist() == [2, 0, 0, 0]


# ---------------------------------------------------------------- dof map --
@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_dof_map_reference_kat(degree):
    tets = np.array([[1, 2, 3, 4], [3, 2, 1, 5]], dtype=np.int32)
    bary, _ = O.lagrange_init(degree)
    st, g, G = O.lagrange_dof_map(degree, tets)
    shared = 0
    for l in range(g.shape[1]):
        if bary[l, 3] != 0:
            continue
        shared += 1
        assert g[0, l] in g[1]
    assert shared == (degree + 1) * (degree + 2) // 2
    assert G == 2 * (degree + 1) * (degree + 2) * (degree + 3) // 6 - shared
    assert O.lagrange_dof_map(0, tets)[0] != 0


@pytest.mark.parametrize("degree,N", [(1, 3), (2, 3), (3, 2), (4, 2)])
def test_dof_map_hash_equals_linear_search(degree, N):
    """The O(N) hash numbering equals the reference's first-appearance linear
    search (tetra_lagrange_global_dof_map.f90:62) restated literally."""
    v, t = O.unit_box_mesh(N)
    s1, g1, c1 = O.lagrange_dof_map(degree, t)
    s2, g2, c2 = O.lagrange_dof_map(degree, t, search=True)
    assert s1 == s2 == 0 and c1 == c2 == (degree * N + 1) ** 3
    assert np.array_equal(g1, g2)
    assert g1[0, 0] == 1  # global dof 1 is the first local dof of cell 1 (its 4th vertex for P1)


def test_dof_map_rejects_repeated_vertices():
    assert O.lagrange_dof_map(1, np.array([[1, 2, 2, 3]], dtype=np.int32))[0] != 0


# ---------------------------------------------------------------- element --
def _poly1d(k, d):
    """coefficients (low -> high) of l_k(t) = prod_{f<k}(d t - f) / k!"""
    c = [Fraction(1)]
    for f in range(k):
        new = [Fraction(0)] * (len(c) + 1)
        for i, a in enumerate(c):
            new[i + 1] += d * a
            new[i] -= f * a
        c = new
    return [a / factorial(k) for a in c]


def _cardinal(bary, d):
    """polynomial dict in (l1, l2, l3, l4) of one cardinal basis function"""
    poly = {(0, 0, 0, 0): Fraction(1)}
    for comp in range(4):
        coeffs = _poly1d(bary[comp], d)
        new = {}
        for mono, a in poly.items():
            for p, b in enumerate(coeffs):
                if b == 0:
                    continue
                m = list(mono)
                m[comp] += p
                new[tuple(m)] = new.get(tuple(m), 0) + a * b
        poly = new
    return poly


def _deriv(poly, comp):
    out = {}
    for mono, a in poly.items():
        if mono[comp] == 0:
            continue
        m = list(mono)
        m[comp] -= 1
        out[tuple(m)] = out.get(tuple(m), 0) + a * mono[comp]
    return out


def _integral(p1, p2):
    """int over the reference tet of p1*p2 (Dirichlet/barycentric formula x 1)"""
    s = Fraction(0)
    for m1, a in p1.items():
        for m2, b in p2.items():
            m = [x + y for x, y in zip(m1, m2)]
            s += a * b * Fraction(np.prod([factorial(x) for x in m]), factorial(sum(m) + 3))
    return s


def exact_element(vertices, degree, kappa, mu):
    bary, _ = O.lagrange_init(degree)
    v = np.asarray(vertices, dtype=float)
    J = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    gl = np.vstack([-Jinv.sum(0), Jinv])  # grad lambda_1..4 (rows)
    polys = [_cardinal(b, degree) for b in bary]
    dpolys = [[_deriv(p, c) for c in range(4)] for p in polys]
    n = len(polys)
    K = np.zeros((n, n))
    for r in range(n):
        for c in range(n):
            s = 0.0
            for c1 in range(4):
                for c2 in range(4):
                    s += float(gl[c1] @ gl[c2]) * float(_integral(dpolys[r][c1], dpolys[c][c2]))
            K[r, c] = det * (kappa * s + mu * float(_integral(polys[r], polys[c])))
    return K


VERTS = np.array([[0.2, -0.3, 0.1], [2.1, -0.5, 0.4], [0.6, 1.4, -0.2], [-0.1, 0.2, 1.6]])


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_element_matches_exact_integration(degree):
    K_exact = exact_element(VERTS, degree, 1.7, 0.6)
    st, K = O.lagrange_stiffness_element(VERTS, degree, 2 * degree + 2, 1.7, 0.6)
    assert st == 0
    assert np.max(np.abs(K - K_exact)) < 1e-12 * np.max(np.abs(K_exact))


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_element_symmetry_reference_kat(degree):
    st, K = O.lagrange_stiffness_element(VERTS, degree, 2 * degree + 2)
    assert st == 0 and np.max(np.abs(K - K.T)) < 2e-11


def test_p1_element_closed_form():
    """P1: K = kappa |T| G G^T with G the barycentric gradients."""
    st, K = O.lagrange_stiffness_element(VERTS, 1, 4, 2.0, 0.0)
    J = np.stack([VERTS[1] - VERTS[0], VERTS[2] - VERTS[0], VERTS[3] - VERTS[0]], axis=1)
    Ji = np.linalg.inv(J)
    grads = np.vstack([-Ji.sum(0), Ji])[[3, 2, 1, 0]]  # local order v4, v3, v2, v1
    vol = np.linalg.det(J) / 6
    assert np.max(np.abs(K - 2.0 * vol * grads @ grads.T)) < 1e-13 * np.max(np.abs(K))


def test_degenerate_element_rejected():
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]])
    assert O.lagrange_stiffness_element(flat, 1, 4)[0] != 0


def test_linear_field_energy_reference_kat():
    mv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1.0]])
    mt = np.array([[1, 2, 3, 4], [1, 3, 2, 5]], dtype=np.int32)
    A, g, G = O.lagrange_stiffness_csc(mv, mt, 1, 4)
    _, nodes = O.lagrange_init(1)
    coef = np.zeros(G)
    for e in range(2):
        vv = mv[mt[e] - 1]
        J = np.stack([vv[1] - vv[0], vv[2] - vv[0], vv[3] - vv[0]], axis=1)
        for l in range(4):
            coef[g[e, l] - 1] = (vv[0] + J @ nodes[l])[0]
    assert abs(coef @ A.matvec(coef) - 1 / 3) < 2e-12


# -------------------------------------------------------------------- CSC --
def test_csc_from_triplet_sums_duplicates():
    rng = np.random.default_rng(3)
    rows = rng.integers(1, 7, 40).astype(np.int32)
    cols = rng.integers(1, 6, 40).astype(np.int32)
    vals = rng.standard_normal(40)
    a = O.csc_from_triplet(6, 5, rows, cols, vals)
    dense = np.zeros((6, 5))
    np.add.at(dense, (rows - 1, cols - 1), vals)
    assert np.allclose(a.to_dense(), dense, atol=1e-15)
    assert a.nnz == len(set(zip(rows.tolist(), cols.tolist())))
    for c in range(5):
        seg = a.row_idx[a.col_ptr[c] - 1:a.col_ptr[c + 1] - 1]
        assert np.all(np.diff(seg) > 0)  # ascending rows, merged


def test_assembled_csc_equals_dense_sum_and_keeps_zeros():
    v, t = O.unit_box_mesh(2)
    A, g, G = O.lagrange_stiffness_csc(v, t, 1, 6)
    dense = np.zeros((G, G))
    for e in range(t.shape[0]):
        st, K = O.lagrange_stiffness_element(v[t[e] - 1], 1, 6)
        dense[np.ix_(g[e] - 1, g[e] - 1)] += K
    assert np.max(np.abs(A.to_dense() - dense)) < 1e-14
    pairs = {(a, b) for e in range(t.shape[0]) for a in g[e] for b in g[e]}
    assert A.nnz == len(pairs)  # structural zeros kept
    assert np.max(np.abs(A.to_dense() - A.to_dense().T)) < 1e-15


# ---------------------------------------------------------------- load/BC --
def test_load_partition_of_unity():
    v, t = O.unit_box_mesh(3)
    for d in (1, 2):
        st, g, G = O.lagrange_dof_map(d, t)
        b = O.lagrange_load(v, t, d, 2 * d + 4, lambda x, y, z: 2.0 + 0 * x, g, G)
        assert abs(b.sum() - 2.0) < 1e-13


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_dirichlet_hash_equals_search_and_boundary(degree):
    v, t = O.unit_box_mesh(3)
    st, g, G = O.lagrange_dof_map(degree, t)
    c1, pts = O.lagrange_dirichlet(v, t, degree, g, G)
    c2, _ = O.lagrange_dirichlet(v, t, degree, g, G, search=True)
    assert np.array_equal(c1, c2)
    # structured closed form: a node is constrained iff it lies on the cube surface
    _, nodes = O.lagrange_init(degree)
    allpts = np.zeros((G, 3))
    for e in range(t.shape[0]):
        vv = v[t[e] - 1]
        J = np.stack([vv[1] - vv[0], vv[2] - vv[0], vv[3] - vv[0]], axis=1)
        allpts[g[e] - 1] = vv[0] + nodes @ J.T
    on = np.any((np.abs(allpts) < 1e-12) | (np.abs(allpts - 1) < 1e-12), axis=1)
    assert np.array_equal(c1, on)
    assert np.allclose(pts[c1], allpts[c1])


def test_constrained_reduction_matches_dense_elimination():
    v, t = O.unit_box_mesh(3)
    A, g, G = O.lagrange_stiffness_csc(v, t, 2, 8, mu=0.3)
    b = np.linspace(-1, 1, G)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    gval = np.where(cons, np.cos(np.arange(G)), 0.0)
    R, free, rb = O.constrained_reduction(A, cons, gval, b)
    D = A.to_dense()
    f = ~cons
    assert np.array_equal(free - 1, np.nonzero(f)[0])
    assert np.max(np.abs(R.to_dense() - D[np.ix_(f, f)])) < 1e-15
    assert np.max(np.abs(rb - (b[f] - D[np.ix_(f, cons)] @ gval[cons]))) < 1e-13


# ----------------------------------------------------------------- Krylov --
def _csc(dense):
    return O.sparse_from_dense(dense)


def test_cg_tridiagonal_reference_kat():
    n = 100
    A = np.diag(np.full(n, 4.1)) - np.eye(n, k=1) - np.eye(n, k=-1)
    x_exact = np.arange(1, n + 1, dtype=float)
    b = A @ x_exact
    x, st = O.solve_sparse(_csc(A), b, method="cg", tolerance=1e-10, max_iterations=n)
    assert st.converged and st.iterations < n / 2
    assert np.linalg.norm(A @ x - b) < 1e-7 and np.linalg.norm(x - x_exact) < 1e-6
    assert st.final_residual < 1e-7
    x, st = O.solve_sparse(_csc(A), b, method="pcg", tolerance=1e-12, max_iterations=5)
    assert not st.converged and st.iterations == 5


def test_cg_distinct_eigenvalue_termination():
    """CG on an SPD matrix with k distinct eigenvalues stops after k steps;
    Jacobi-PCG on a diagonal matrix stops after one."""
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    lam = np.repeat([1.0, 2.0, 5.0, 9.0], [8, 7, 8, 7])
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    b = rng.standard_normal(30)
    x, st = O.solve_sparse(_csc(A), b, method="cg", preconditioner="none", tolerance=1e-10,
                           max_iterations=30)
    assert st.converged and st.iterations == 4
    assert np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-8 * np.linalg.norm(x)
    D = np.diag(rng.uniform(1, 100, 25))
    x, st = O.solve_sparse(_csc(D), np.ones(25), method="pcg", preconditioner="jacobi", tolerance=1e-12,
                           max_iterations=10)
    assert st.iterations == 1 and np.allclose(x, 1 / np.diag(D), rtol=1e-14)


def test_pcg_jacobi_against_direct_solve():
    v, t = O.unit_box_mesh(4)
    A, g, G = O.lagrange_stiffness_csc(v, t, 2, 8)
    b = O.lagrange_load(v, t, 2, 8, lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z), g, G)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    R, free, rb = O.constrained_reduction(A, cons, np.zeros(G), b)
    x, st = O.solve_sparse(R, rb, tolerance=1e-12, max_iterations=2000, preconditioner="jacobi",
                           breakdown="relative")
    xd = np.linalg.solve(R.to_dense(), rb)
    assert st.converged and np.linalg.norm(x - xd) < 1e-10 * np.linalg.norm(xd)
    # the reference's absolute breakdown test stops this small-h solve early
    x0, st0 = O.solve_sparse(R, rb, tolerance=1e-12, max_iterations=2000, preconditioner="jacobi")
    assert not st0.converged and st0.iterations < st.iterations


def test_quartic_solution_reproduced_at_degree_four():
    """test_tetra_lagrange_poisson_p_convergence.f90:26-40: u = xyz(1-x-y-z),
    f = -lap u = 2(xy + xz + yz), u = 0 on the boundary of the unit tet split
    around its centroid; P4 reproduces u at every node."""
    mv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.25, 0.25, 0.25]])
    mt = np.array([[1, 2, 3, 5], [1, 4, 2, 5], [1, 3, 4, 5], [2, 4, 3, 5]], dtype=np.int32)
    d, q = 4, 12
    A, g, G = O.lagrange_stiffness_csc(mv, mt, d, q)
    b = O.lagrange_load(mv, mt, d, q, lambda x, y, z: 2 * (x * y + x * z + y * z), g, G)
    cons, pts = O.lagrange_dirichlet(mv, mt, d, g, G)
    R, free, rb = O.constrained_reduction(A, cons, np.zeros(G), b)
    x = np.linalg.solve(R.to_dense(), rb)
    _, nodes = O.lagrange_init(d)
    node_pts = np.zeros((G, 3))
    for e in range(4):
        vv = mv[mt[e] - 1]
        J = np.stack([vv[1] - vv[0], vv[2] - vv[0], vv[3] - vv[0]], axis=1)
        node_pts[g[e] - 1] = vv[0] + nodes @ J.T
    P = node_pts[free - 1]
    u = P[:, 0] * P[:, 1] * P[:, 2] * (1 - P.sum(1))
    assert np.max(np.abs(x - u)) < 1e-12
