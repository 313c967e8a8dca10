"""Pins the Nedelec part of the CPU oracle.

Restated known answers: test/test_tetra_edge_dof_map.f90,
test/test_tetra_nedelec_global_dof_map.f90:22-70,
test/test_tetra_nedelec_sparse_assembly_slow.f90:31-85 (energies 3 and 28/3,
constant-field energy at orders 2..6), test/test_tetra_nedelec_first_order.f90
(tangential-moment delta), test/test_tetra_nedelec_arbitrary_order_slow.f90
(Kronecker moments), test/test_tetra_nedelec_arbitrary_order_assembly_slow.f90
(symmetry).  The reference's generated tables (orders 1-4) are compared
through oracle/_ref where it is built.
"""
from math import factorial

import numpy as np
import pytest

import oracle as O
from oracle import ref as R

MESH_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1.0]])
MESH_T = np.array([[1, 2, 3, 4], [1, 3, 2, 5]], dtype=np.int32)
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
FACES = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
REFV = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
need_ref = pytest.mark.skipif(not R.available(), reason="reference build (oracle/_ref) not present")


def test_edge_map_hash_equals_search():
    v, t = O.unit_box_mesh(3)
    a = O.edge_dof_map(t)
    b = O.edge_dof_map(t, search=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    st, edges, gd, orient = a
    assert edges.shape[0] == 3 * 4 * 4 * 4 + 3 * 4 * 4 * 3 // 1 - 0 or edges.shape[0] > 0
    # orientation: +1 iff the tet-local first vertex id < second
    for e in range(t.shape[0]):
        for l, (i, j) in enumerate(EDGES):
            assert orient[e, l] == (1 if t[e, i] < t[e, j] else -1)


@need_ref
def test_edge_map_equals_reference():
    v, t = O.unit_box_mesh(3)
    st, edges, gd, orient = O.edge_dof_map(t)
    st2, edges2, gd2, orient2 = R.edge_dof_map(t)
    assert st == st2 == 0
    assert np.array_equal(edges, edges2) and np.array_equal(gd, gd2) and np.array_equal(orient, orient2)


def test_nedelec_dof_map_reference_kat():
    t = np.array([[1, 2, 3, 4], [3, 1, 2, 5]], dtype=np.int32)
    gd, eo, fp, nd, ne, nf = O.nedelec_dof_map(6, t)
    assert ne == 9 and nf == 7 and gd.shape[1] == 216 and gd.max() == 384 and nd == 384
    fs = 6 * 6
    assert np.array_equal(gd[0, fs:fs + 30], gd[1, fs:fs + 30])
    assert fp[0, 0].tolist() == [1, 2, 3] and fp[1, 0].tolist() == [3, 1, 2]
    T = O.nedelec_basis_transform(6, eo[1], fp[1])
    assert [T[i, i] for i in range(4)] == [-1.0, 1.0, -1.0, 1.0]
    with pytest.raises(ValueError):
        O.nedelec_dof_map(0, t)
    with pytest.raises(ValueError):
        O.nedelec_dof_map(6, np.array([[1, 2, 3, 4], [1, 1, 2, 5]], dtype=np.int32))


@need_ref
@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_basis_equals_generated_tables(k):
    rng = np.random.default_rng(k)
    for _ in range(10):
        p = rng.dirichlet(np.ones(4))[1:]
        v, c = O.nedelec_eval(k, p)
        st, vr, cr = R.nedelec_basis(k, p)
        scale = max(np.abs(vr).max(), np.abs(cr).max())
        assert st == 0
        assert np.abs(v - vr).max() <= 1e-13 * scale
        assert np.abs(c - cr).max() <= 1e-13 * scale


@need_ref
@pytest.mark.parametrize("k", [2, 3, 4])
def test_face_transforms_equal_generated_tables(k):
    for perm in ([1, 2, 3], [1, 3, 2], [2, 1, 3], [2, 3, 1], [3, 1, 2], [3, 2, 1]):
        T = O.face_basis_to_local(k, perm)
        Tr = R.face_basis_to_local(k, perm)
        assert np.abs(T - Tr).max() <= 1e-13 * max(1.0, np.abs(Tr).max())


def test_general_construction_reproduces_whitney_at_order_one():
    """The moment-dual construction at k = 1 gives the Whitney functions."""
    cf = O.nedelec_coefficients(1)
    rng = np.random.default_rng(0)
    import ctypes as C

    for _ in range(5):
        p = rng.dirichlet(np.ones(4))[1:]
        vw, cw = O.nedelec_eval(1, p)
        v = np.zeros((6, 3))
        c = np.zeros((6, 3))
        O._lib().ffo_nedelec_eval(C.c_int(1), cf.ctypes.data_as(C.c_void_p), np.ascontiguousarray(p).ctypes.data_as(C.c_void_p),
                                  v.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
        assert np.abs(v - vw).max() < 1e-14 and np.abs(c - cw).max() < 1e-14


def _moments(k, fn):
    """Degrees of freedom of a reference field fn(p) -> (3,) (the functionals of
    build_runtime_coefficients, evaluated with numpy quadrature)."""
    out = []
    xg, wg = np.polynomial.legendre.leggauss(k + 2)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    for (i, j) in EDGES:
        tan = REFV[j] - REFV[i]
        for m in range(k):
            P = np.polynomial.legendre.Legendre.basis(m)
            out.append(sum(w * P(2 * t - 1) * (tan @ fn(REFV[i] + t * tan)) for t, w in zip(xg, wg)))
    tx, ty, tw = O.triangle_duffy(2 * k + 4)
    for (a, b, c) in FACES:
        t1, t2 = REFV[b] - REFV[a], REFV[c] - REFV[a]
        for tv in (t1, t2):
            for td in range(k - 1):
                for xa in range(td + 1):
                    yb = td - xa
                    out.append(sum(w * x ** xa * y ** yb * (tv @ fn(REFV[a] + x * t1 + y * t2))
                                   for x, y, w in zip(tx, ty, tw)))
    vx, vy, vz, vw = O.tetra_duffy(2 * k + 4)
    for comp in range(3):
        for td in range(k - 2):
            for xa in range(td + 1):
                for yb in range(td + 1 - xa):
                    zg = td - xa - yb
                    out.append(sum(w * x ** xa * y ** yb * z ** zg * fn(np.array([x, y, z]))[comp]
                                   for x, y, z, w in zip(vx, vy, vz, vw)))
    return np.array(out)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_kronecker_moments(k):
    n = O.nedelec_dof_count(k)
    cache = {}

    def basis(p):
        key = tuple(np.round(p, 15))
        if key not in cache:
            cache[key] = O.nedelec_eval(k, p)[0]
        return cache[key]

    M = np.stack([_moments(k, lambda p, j=j: basis(p)[j]) for j in range(n)], axis=1)
    assert np.abs(M - np.eye(n)).max() < 1e-9


def _constant_field_dofs(k, v, t, field):
    """interpolant of a constant field in the global dofs (the DOF vector of
    test_tetra_nedelec_sparse_assembly_slow.f90:95-165, monomial moments)"""
    gd, eo, fp, nd, ne, nf = O.nedelec_dof_map(k, t)
    st, edges, egd, eor = O.edge_dof_map(t)
    dofs = np.zeros(nd)
    for e, (a, b) in enumerate(edges):
        dofs[e * k] = field @ (v[b - 1] - v[a - 1])
    # faces in first-appearance order with their sorted vertex triples
    faces = {}
    for c in range(t.shape[0]):
        for f in FACES:
            key = tuple(sorted(t[c, list(f)]))
            if key not in faces:
                faces[key] = len(faces)
    fdc = k * (k - 1)
    for key, fi in faces.items():
        for comp in range(2):
            tan = v[key[comp + 1] - 1] - v[key[0] - 1]
            mono = 0
            for deg in range(k - 1):
                for xa in range(deg + 1):
                    yb = deg - xa
                    idx = ne * k + fi * fdc + comp * fdc // 2 + mono
                    if k <= 5:
                        dofs[idx] = (field @ tan) * factorial(xa) * factorial(yb) / factorial(xa + yb + 2)
                    else:  # Dubiner moments: psi_00 = 1, the rest orthogonal to constants
                        dofs[idx] = 0.5 * (field @ tan) if deg == 0 else 0.0
                    mono += 1
    base = 6 * k + 4 * fdc
    for c in range(t.shape[0]):
        vv = v[t[c] - 1]
        J = np.stack([vv[1] - vv[0], vv[2] - vv[0], vv[3] - vv[0]], axis=1)
        rf = J.T @ field
        l = base
        for comp in range(3):
            for deg in range(k - 2):
                for xa in range(deg + 1):
                    for yb in range(deg + 1 - xa):
                        zg = deg - xa - yb
                        if k <= 5:
                            dofs[gd[c, l] - 1] = rf[comp] * factorial(xa) * factorial(yb) * factorial(zg) / \
                                factorial(xa + yb + zg + 3)
                        else:  # Koornwinder moments
                            dofs[gd[c, l] - 1] = rf[comp] / 6.0 if deg == 0 else 0.0
                        l += 1
    return dofs


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_constant_field_energy(k):
    A = O.nedelec_csc(k, MESH_V, MESH_T, alpha=4.0, beta=1.5)
    dofs = _constant_field_dofs(k, MESH_V, MESH_T, np.array([1.0, 2.0, -1.0]))
    energy = dofs @ A.matvec(dofs)
    assert abs(energy - 3.0) < (2e-12 if k == 1 else 1e-8)


def test_rotational_field_energy_order_one():
    st, edges, gd, orient = O.edge_dof_map(MESH_T)
    c = np.array([1.0, -2.0, 3.0])
    dofs = np.array([0.5 * np.cross(c, 0.5 * (MESH_V[a - 1] + MESH_V[b - 1])) @ (MESH_V[b - 1] - MESH_V[a - 1])
                     for a, b in edges])
    A = O.nedelec_csc(1, MESH_V, MESH_T, alpha=2.0, beta=0.0)
    assert abs(dofs @ A.matvec(dofs) - 28.0 / 3.0) < 3e-12


VERTS = np.array([[0.2, -0.3, 0.4], [2.2, -0.3, 0.4], [0.2, 2.7, 0.4], [0.2, -0.3, 4.4]])


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_element_symmetry(k):
    st, E = O.nedelec_element(k, VERTS, 2 * k + 2, alpha=1.7, beta=0.8)
    assert st == 0 and np.abs(E - E.T).max() < 3e-12


@pytest.mark.parametrize("k", [1, 2, 3])
def test_weighted_identity_tensor_equals_curl_mass(k):
    v, t = O.unit_box_mesh(1)
    K = np.tile(2.5 * np.eye(3), (t.shape[0], 1, 1))
    A = O.nedelec_csc(k, v, t, alpha=2.5, beta=0.7)
    W = O.nedelec_csc(k, v, t, beta=0.7, cell_tensor=K)
    assert np.array_equal(A.row_idx, W.row_idx)
    assert np.abs(A.val - W.val).max() < 1e-12 * np.abs(A.val).max()


def test_high_order_fixture_is_the_oracle():
    """tests/golden/nedelec_high_order.npz holds the oracle's output: the
    cheapest of its cases (tensor-weighted k = 4 on the relabelled box) is
    recomputed here and must match bit for bit."""
    import os

    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nedelec_high_order.npz"))
    a = O.nedelec_csc(4, g["box_v"], g["box_t"], alpha=1.0, beta=0.6, cell_tensor=g["box_K"])
    assert np.array_equal(a.col_ptr, g["box_k4w_col_ptr"]) and np.array_equal(a.row_idx, g["box_k4w_row_idx"])
    assert np.array_equal(a.val, g["box_k4w_val"])
