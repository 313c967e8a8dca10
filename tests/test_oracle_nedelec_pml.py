"""The oracle's complex PML Nedelec assembly (ffo_nedelec_pml_triplets).

assemble_tetra_nedelec_pml_csc (src/assembly/assembly_tetra_nedelec_3d.f90:
1613-1685) and assemble_tetra_nedelec_curvilinear_pml_csc (:1372-1444), with
the coefficients of src/boundary/cartesian_helmholtz_pml.f90 and
curvilinear_helmholtz_pml.f90:113-131.

Pinned by: the reference's constant-field energy on its two-tet mesh
(test/test_tetra_nedelec_pml_sparse.f90:20-51, 3e-12); the closed-form energy
of a linear field a + b x x with complex stretches (constant curl b2 = 2b:
vol * b2.K.b2, and the exact quadratic integral of each field component);
and coordinate-change identities that hold for a real stretch S: the PML
operator equals the curl-mass operator (alpha 1, beta -k^2) on the mesh
mapped by x -> S x, cartesian (S diagonal) and curvilinear alike.
"""
import numpy as np
import pytest

import oracle as O

REF_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
REF_T = np.array([[1, 2, 3, 4], [1, 3, 2, 5]], dtype=np.int32)


def dense(a):
    """Square CSC (oracle or product type) to a dense array."""
    n = a.col_ptr.shape[0] - 1
    A = np.zeros((n, n), dtype=a.val.dtype)
    for c in range(n):
        for k in range(a.col_ptr[c] - 1, a.col_ptr[c + 1] - 1):
            A[a.row_idx[k] - 1, c] += a.val[k]
    return A


def masses(s):
    return np.array([s[1] * s[2] / s[0], s[0] * s[2] / s[1], s[0] * s[1] / s[2]])


def whitney_dofs(v, t, field):
    """Order-1 edge dofs of a field linear in x: its tangential integral along
    each global edge (exact at the midpoint)."""
    _, edges, _, _ = O.edge_dof_map(t)
    out = []
    for e in edges:
        a, b = v[e[0] - 1], v[e[1] - 1]
        out.append(field(0.5 * (a + b)) @ (b - a))
    return np.array(out)


def test_reference_constant_field_energy():
    s = np.array([[1 + 0.6j, 1, 1]] * 2)
    a = O.nedelec_pml_csc(1, REF_V, REF_T, s, 1.3)
    f = np.array([1.0, 2.0, -1.0])
    d = whitney_dofs(REF_V, REF_T, lambda x: f)
    energy = d @ (dense(a) @ d)
    expected = -1.3 ** 2 / 3.0 * np.sum(masses(s[0]) * f ** 2)
    assert abs(energy - expected) < 3.0e-12
    for k in (2, 3, 4):
        assert O.nedelec_pml_csc(k, REF_V, REF_T, s, 1.3).n_rows > 0


def test_linear_field_closed_form_complex_stretch():
    v, t = O.unit_box_mesh(2)
    s0 = np.array([1.0 + 0.4j, 0.8 - 0.3j, 1.3 + 0.2j])
    s = np.broadcast_to(s0, (t.shape[0], 3))
    k0 = 0.9
    a0 = np.array([0.3, -0.5, 0.7])
    b = np.array([0.4, 0.9, -0.6])
    field = lambda x: a0 + np.cross(b, x)
    d = whitney_dofs(v, t, field)
    energy = d @ (dense(O.nedelec_pml_csc(1, v, t, s, k0)) @ d)
    m = masses(s0)
    curl = 2.0 * b
    vol = 0.0
    quad = np.zeros(3)
    for c in t:
        x = v[c - 1]
        det = abs(np.linalg.det(np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1))) / 6.0
        vol += det
        p = np.array([field(xa) for xa in x])  # (4, 3)
        quad += det / 20.0 * (np.sum(p ** 2, axis=0) + np.sum(p, axis=0) ** 2)
    expected = vol * np.sum(curl ** 2 / m) - k0 ** 2 * np.sum(m * quad)
    assert abs(energy - expected) <= 1e-12 * abs(expected)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_real_cartesian_stretch_is_a_scaled_mesh(k):
    v, t = O.unit_box_mesh(2)
    s = np.array([1.5, 0.7, 1.2])
    a = O.nedelec_pml_csc(k, v, t, np.broadcast_to(s, (t.shape[0], 3)), 1.3)
    c = O.nedelec_csc(k, v * s, t, alpha=1.0, beta=-1.3 ** 2)
    assert np.array_equal(a.row_idx, c.row_idx) and np.array_equal(a.col_ptr, c.col_ptr)
    assert np.max(np.abs(a.val - c.val)) <= 1e-13 * np.max(np.abs(c.val))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_real_curvilinear_stretch_is_a_mapped_mesh(k):
    v, t = O.unit_box_mesh(2)
    S = np.eye(3) + np.random.default_rng(3).uniform(-0.3, 0.3, (3, 3))
    assert np.linalg.det(S) > 0.1
    a = O.nedelec_pml_csc(k, v, t, np.broadcast_to(S, (t.shape[0], 3, 3)), 0.9, kind=1)
    c = O.nedelec_csc(k, v @ S.T, t, alpha=1.0, beta=-0.81)
    assert np.array_equal(a.row_idx, c.row_idx)
    assert np.max(np.abs(a.val - c.val)) <= 1e-13 * np.max(np.abs(c.val))


def test_diagonal_curvilinear_equals_cartesian():
    v, t = O.unit_box_mesh(2)
    rng = np.random.default_rng(4)
    s = rng.uniform(0.5, 1.5, (t.shape[0], 3)) + 1j * rng.uniform(-0.5, 0.5, (t.shape[0], 3))
    D = np.zeros((t.shape[0], 3, 3), dtype=complex)
    for i in range(3):
        D[:, i, i] = s[:, i]
    a1 = O.nedelec_pml_csc(2, v, t, s, 1.1)
    a2 = O.nedelec_pml_csc(2, v, t, D, 1.1, kind=1)
    assert np.max(np.abs(a1.val - a2.val)) <= 1e-13 * np.max(np.abs(a1.val))


def test_rejects_zero_or_singular_stretch():
    s = np.array([[1.0, 0.0, 1.0]] * 2, dtype=complex)
    with pytest.raises(ValueError):
        O.nedelec_pml_csc(1, REF_V, REF_T, s, 1.3)
    S = np.zeros((2, 3, 3), dtype=complex)
    S[:, 0, 0] = S[:, 1, 1] = 1.0
    with pytest.raises(ValueError):
        O.nedelec_pml_csc(1, REF_V, REF_T, S, 1.3, kind=1)
