"""Oracle restatement of the H1 assembly tangents, pinned on the CPU.

assemble_tetra_lagrange_stiffness_csc_jvp / _vjp
(src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:107-275,
336-471).  The two-tet fixture, step and bars are the reference's own
(test/test_tetra_lagrange_stiffness_csc_ad.f90); the other checks hold by
the mathematics alone: the stiffness is invariant under translations and
rotations of the mesh and scales like h (the mass like h^3) under
dilation, and the matrix is linear in both coefficients, so
kappa_bar = <values_bar, K(kappa=1, mu=0)> and mu_bar = <values_bar, M>.
A dropped determinant tangent, a transposed inverse or a sign slip in any
chain step fails one of them.
"""
import numpy as np
import pytest

import oracle as O

# test/test_tetra_lagrange_stiffness_csc_ad.f90:10-39
FIX_TETS = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], dtype=np.int32)
FIX_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
FIX_VD = np.array([[0.01, -0.02, 0.015], [-0.03, 0.01, 0.02], [0.025, -0.015, 0.005],
                   [-0.01, 0.03, -0.02], [0.02, -0.01, 0.025]])
FIX_K, FIX_M, FIX_KD, FIX_MD = 1.7, 0.8, 0.13, -0.09


def fixture_values_bar(nnz):
    e = np.arange(1, nnz + 1, dtype=np.float64)
    return 0.003 * e - 0.00001 * e ** 2


def distorted_box(N, seed):
    v, t = O.unit_box_mesh(N)
    rng = np.random.default_rng(seed)
    return v + (0.25 / N) * rng.uniform(-1, 1, v.shape), t


def test_reference_fixture():
    d, q, step = 2, 6, 2.0e-7
    a, _, _ = O.lagrange_stiffness_csc(FIX_V, FIX_TETS, d, q, FIX_K, FIX_M)
    ad = O.lagrange_stiffness_csc_jvp(FIX_V, FIX_TETS, d, q, FIX_K, FIX_M, FIX_VD, FIX_KD, FIX_MD)
    ap, _, _ = O.lagrange_stiffness_csc(FIX_V + step * FIX_VD, FIX_TETS, d, q, FIX_K + step * FIX_KD,
                                        FIX_M + step * FIX_MD)
    am, _, _ = O.lagrange_stiffness_csc(FIX_V - step * FIX_VD, FIX_TETS, d, q, FIX_K - step * FIX_KD,
                                        FIX_M - step * FIX_MD)
    assert ad.nnz == a.nnz == ap.nnz == am.nnz
    assert np.array_equal(ad.col_ptr, a.col_ptr) and np.array_equal(ad.row_idx, a.row_idx)
    err = np.max(np.abs(ad.val - (ap.val - am.val) / (2 * step))) / max(1.0, np.max(np.abs(ad.val)))
    assert err < 5.0e-8
    vb_in = fixture_values_bar(a.nnz)
    vb, kb, mb, _ = O.lagrange_stiffness_csc_vjp(FIX_V, FIX_TETS, d, q, vb_in, FIX_K, FIX_M)
    lhs = vb_in @ ad.val
    rhs = np.sum(vb * FIX_VD) + kb * FIX_KD + mb * FIX_MD
    assert abs(lhs - rhs) < 3.0e-10 * max(1.0, abs(lhs), abs(rhs))


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_rigid_motions_and_dilation(d):
    v, t = distorted_box(2, d)
    q = 2 * d
    K, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=1.0, mu=0.0)
    M, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=0.0, mu=1.0)
    scale = np.max(np.abs(K.val))
    # translation: zero tangent
    tr = np.tile([0.3, -0.7, 1.1], (v.shape[0], 1))
    assert np.max(np.abs(O.lagrange_stiffness_csc_jvp(v, t, d, q, 1.3, 0.6, tr).val)) <= 1e-12 * scale
    # rotation generator x -> w x x: zero tangent
    w = np.array([0.2, -0.5, 0.9])
    rot = np.cross(w, v)
    assert np.max(np.abs(O.lagrange_stiffness_csc_jvp(v, t, d, q, 1.3, 0.6, rot).val)) <= 1e-12 * scale
    # dilation x -> (1 + s) x: dK = K, dM = 3 M
    dil = O.lagrange_stiffness_csc_jvp(v, t, d, q, 1.3, 0.6, v)
    assert np.max(np.abs(dil.val - (1.3 * K.val + 3 * 0.6 * M.val))) <= 1e-12 * scale


@pytest.mark.parametrize("d", [1, 3])
def test_vjp_invariants_and_coefficient_bars(d):
    v, t = distorted_box(2, 10 + d)
    q = 2 * d
    K, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=1.0, mu=0.0)
    M, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=0.0, mu=1.0)
    rng = np.random.default_rng(d)
    vb_in = rng.uniform(-1, 1, K.nnz)
    vb, kb, mb, ckb = O.lagrange_stiffness_csc_vjp(v, t, d, q, vb_in, 1.3, 0.6)
    ref = max(1.0, np.max(np.abs(vb)))
    # translation and rotation invariance of K + M: the vertex cotangent has
    # zero net force and zero net torque
    assert np.max(np.abs(vb.sum(axis=0))) <= 1e-12 * ref * v.shape[0]
    assert np.max(np.abs(np.cross(v, vb).sum(axis=0))) <= 1e-12 * ref * v.shape[0]
    # linearity in the coefficients
    assert abs(kb - vb_in @ K.val) <= 1e-12 * max(1.0, abs(kb))
    assert abs(mb - vb_in @ M.val) <= 1e-12 * max(1.0, abs(mb))
    assert abs(ckb.sum() - kb) <= 1e-12 * max(1.0, abs(kb))


def test_cell_kappa_tangent_matches_difference():
    v, t = distorted_box(2, 20)
    d, q = 2, 4
    rng = np.random.default_rng(21)
    ck = np.exp(rng.normal(size=t.shape[0]))
    ckd = rng.normal(size=t.shape[0])
    ad = O.lagrange_stiffness_csc_jvp(v, t, d, q, 1.0, 0.4, None, 0.0, 0.0, cell_kappa=ck,
                                      cell_kappa_dot=ckd)
    # linear in kappa: the tangent is exactly the matrix with kappa_dot
    ak, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=1.0, mu=0.0, cell_kappa=ckd)
    assert np.max(np.abs(ad.val - ak.val)) <= 1e-13 * np.max(np.abs(ak.val))
    vb_in = rng.uniform(-1, 1, ad.nnz)
    _, kb, _, ckb = O.lagrange_stiffness_csc_vjp(v, t, d, q, vb_in, 1.0, 0.4, cell_kappa=ck)
    assert abs(ckb @ ckd - vb_in @ ad.val) <= 1e-12 * max(1.0, abs(vb_in @ ad.val))
