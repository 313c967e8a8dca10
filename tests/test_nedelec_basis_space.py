"""Checks of the library's Nedelec basis that do not depend on how the face
and interior moments are normalised (host routine ffh_nedelec_basis, no GPU).

The first-kind Nedelec space of order k on a tetrahedron is
(P_{k-1})^3 + S_k with S_k the homogeneous degree-k fields p with p . x = 0;
it contains every field of degree <= k - 1 and the gradient of every
polynomial of degree <= k (test/test_tetra_nedelec_arbitrary_order_slow.f90:
25-83 checks the same properties through the moments).  At order 6 the
reference's moments use fortnum's Dubiner / Koornwinder polynomials, whose
normalisation is not available, so entry-wise parity at k = 6 is unpinned;
these checks pin the space itself: the n = k (k + 2) (k + 3) / 2 functions
are independent, span (P_{k-1})^3 (so every gradient of a degree-k
polynomial) and the degree-k fields of S_k exactly, do not reach a degree-k
field outside the space, and their curls are the curls of their values.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


@pytest.mark.parametrize("k", [5, 6])
def test_basis_spans_first_kind_space(ff, k):
    n = k * (k + 2) * (k + 3) // 2
    rng = np.random.default_rng(k)
    pts = rng.dirichlet(np.ones(4), size=4 * n)[:, 1:]
    vals, curls = ff.evaluate_tetra_nedelec_first_kind(k, pts)   # (npts, n, 3)
    B = vals.transpose(0, 2, 1).reshape(-1, n)                    # rows (point, component)
    sv = np.linalg.svd(B, compute_uv=False)
    assert sv[-1] > 1e-10 * sv[0]                                 # n independent functions
    x, y, z = pts.T

    def residual(field):
        f = field.reshape(-1)
        c, *_ = np.linalg.lstsq(B, f, rcond=None)
        return np.linalg.norm(B @ c - f) / np.linalg.norm(f)

    for (a, b, c) in [(k - 1, 0, 0), (0, 2, k - 3), (1, 1, k - 3), (2, k - 4, 1)]:
        m = x ** a * y ** b * z ** c
        for comp in range(3):
            f = np.zeros((pts.shape[0], 3))
            f[:, comp] = m
            assert residual(f) < 1e-9
    for (a, b, c) in [(k, 0, 0), (2, 3, k - 5), (0, 1, k - 1), (3, k - 3, 0)]:
        # gradients of degree-k polynomials
        g = np.stack([a * x ** max(a - 1, 0) * y ** b * z ** c,
                      b * x ** a * y ** max(b - 1, 0) * z ** c,
                      c * x ** a * y ** b * z ** max(c - 1, 0)], 1)
        assert residual(g) < 1e-9
    # degree-k fields of S_k (homogeneous, p . x = 0): x^a y^b z^c (y, -x, 0)
    # and x^a y^b z^c (0, z, -y) with a + b + c = k - 1
    for (a, b, c) in [(k - 1, 0, 0), (1, k - 3, 1), (0, 0, k - 1)]:
        m = x ** a * y ** b * z ** c
        assert residual(np.stack([m * y, -m * x, 0 * m], 1)) < 1e-9
        assert residual(np.stack([0 * m, m * z, -m * y], 1)) < 1e-9
    # a field of degree k outside the space: x^k e_1 is not a gradient and
    # not in S_k (x . (x^k, 0, 0) != 0), so it is not reproduced
    f = np.zeros((pts.shape[0], 3))
    f[:, 0] = x ** k
    assert residual(f) > 1e-3
    # curls are the curls of the values (central differences)
    h = 1e-5
    p0 = pts[:16]
    d = []
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = h
        vp, _ = ff.evaluate_tetra_nedelec_first_kind(k, p0 + e)
        vm, _ = ff.evaluate_tetra_nedelec_first_kind(k, p0 - e)
        d.append((vp - vm) / (2 * h))                             # d/dx_ax, (16, n, 3)
    cd = np.stack([d[1][..., 2] - d[2][..., 1], d[2][..., 0] - d[0][..., 2],
                   d[0][..., 1] - d[1][..., 0]], -1)
    _, c0 = ff.evaluate_tetra_nedelec_first_kind(k, p0)
    assert np.max(np.abs(cd - c0)) <= 1e-5 * np.max(np.abs(c0))
