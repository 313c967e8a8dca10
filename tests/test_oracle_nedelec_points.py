"""The oracle's weighted Nedelec element with the coefficient evaluated at
every point of the rule (assemble_tetra_nedelec_weighted_element(_order),
assembly_tetra_nedelec_3d.f90:2406-2514, coefficient(x_q) at :2435 / :2499),
pinned by what the mathematics fixes:

* tensors equal at all points reproduce the cell-constant weighted element
  (itself pinned against fixtures and energies in test_oracle_nedelec.py);
* at k = 1 the curls are constant on a cell and the 4-point rule integrates
  linear functions exactly, so a tensor field linear in x gives the element
  of the constant tensor K(centroid);
* the curl-curl form is invariant under translating the cell and the field
  together, and under rotating both (K'(x) = R K(R^T x) R^T): a point map
  that drops v1 or uses J^T instead of J fails these;
* a field quadratic in x is integrated exactly by the rule, so the element
  does not change when the rule is refined by four degrees.
"""
import numpy as np
import pytest

import oracle as O

V4 = np.array([[0.1, -0.2, 0.05], [1.3, 0.1, -0.1], [0.2, 0.9, 0.3], [-0.1, 0.25, 1.1]])


def sym(a):
    return 0.5 * (a + a.T)


def field(c0, c1, c2, quad=False):
    """K(x) = K0 + x0 K1 + x1 K2 (+ x2^2 K1 when quad), symmetric."""
    def K(x):
        k = c0 + x[0] * c1 + x[1] * c2
        if quad:
            k = k + x[2] ** 2 * c1
        return k
    return K


@pytest.fixture(scope="module")
def tensors():
    rng = np.random.default_rng(7)
    c0 = np.eye(3) * 3 + sym(rng.normal(size=(3, 3))) * 0.3
    c1 = sym(rng.normal(size=(3, 3)))
    c2 = sym(rng.normal(size=(3, 3)))
    return c0, c1, c2


def element_of_field(k, v4, K, beta=0.7, qdeg=None):
    pts = O.nedelec_rule_points(k, qdeg)
    J = np.stack([v4[1] - v4[0], v4[2] - v4[0], v4[3] - v4[0]], axis=1)
    X = v4[0] + pts[:, :3] @ J.T
    Kp = np.array([K(x) for x in X])
    st, m = O.nedelec_element_points(k, v4, Kp, beta=beta, qdeg=qdeg)
    assert st == 0
    return m


@pytest.mark.parametrize("k", [1, 2, 3])
def test_constant_points_reproduce_cell_tensor_csc(k, tensors):
    v, t = O.box_mesh([[0, 0, 0], [1, 1.2, 0.9]], [2, 1, 1])
    rng = np.random.default_rng(k)
    Kc = np.array([sym(rng.normal(size=(3, 3))) + 3 * np.eye(3) for _ in range(t.shape[0])])
    nq = O.nedelec_rule_points(k).shape[0]
    Kp = np.repeat(Kc[:, None], nq, axis=1)
    a = O.nedelec_csc(k, v, t, beta=0.8, cell_tensor=Kc)
    b = O.nedelec_csc_points(k, v, t, Kp, beta=0.8)
    assert np.array_equal(a.col_ptr, b.col_ptr) and np.array_equal(a.row_idx, b.row_idx)
    assert np.max(np.abs(a.val - b.val)) <= 1e-15 * np.max(np.abs(a.val))


def test_k1_linear_field_equals_centroid_tensor(tensors):
    c0, c1, c2 = tensors
    K = field(c0, c1, c2)
    m = element_of_field(1, V4, K)
    st, mc = O.nedelec_element(1, V4, alpha=1.0, beta=0.7, tensor=K(V4.mean(axis=0)))
    assert st == 0
    assert np.max(np.abs(m - mc)) <= 1e-13 * np.max(np.abs(mc))


@pytest.mark.parametrize("k", [2, 3])
def test_translation_and_rotation_invariance(k, tensors):
    c0, c1, c2 = tensors
    K = field(c0, c1, c2)
    m = element_of_field(k, V4, K)
    shift = np.array([0.7, -1.3, 2.1])
    mt = element_of_field(k, V4 + shift, lambda x: K(x - shift))
    assert np.max(np.abs(mt - m)) <= 1e-12 * np.max(np.abs(m))
    th = 0.63
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    R = R @ np.array([[1, 0, 0], [0, np.cos(0.4), -np.sin(0.4)], [0, np.sin(0.4), np.cos(0.4)]])
    mr = element_of_field(k, V4 @ R.T, lambda x: R @ K(R.T @ x) @ R.T)
    assert np.max(np.abs(mr - m)) <= 1e-12 * np.max(np.abs(m))
    # a point map without the v1 offset breaks translation invariance
    pts = O.nedelec_rule_points(k)
    J = np.stack([V4[1] - V4[0], V4[2] - V4[0], V4[3] - V4[0]], axis=1)
    wrong = np.array([K(J @ p) for p in pts[:, :3]])
    st, mw = O.nedelec_element_points(k, V4, wrong, beta=0.7)
    assert np.max(np.abs(mw - m)) > 1e-6 * np.max(np.abs(m))


@pytest.mark.parametrize("k", [2, 3])
def test_quadratic_field_is_integrated_exactly(k, tensors):
    c0, c1, c2 = tensors
    K = field(c0, c1, c2, quad=True)
    m = element_of_field(k, V4, K)
    mf = element_of_field(k, V4, K, qdeg=2 * k + 6)
    assert np.max(np.abs(mf - m)) <= 1e-12 * np.max(np.abs(m))


def test_rule_points_are_the_load_rule():
    """The weighted element and the vector load share one rule: Duffy of
    degree 2k + 2 (:1977 passes it to the element, :546-547 for the load)
    and the 4-point rule at k = 1 (:2424, :2533)."""
    import fortfem_amd as ff

    for k in (1, 2, 4):
        assert np.allclose(O.nedelec_rule_points(k), ff.nedelec_load_points(k), rtol=0, atol=1e-15)
