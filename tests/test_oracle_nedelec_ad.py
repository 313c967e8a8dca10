"""Oracle restatement of the Nedelec curl-mass assembly tangents, pinned on
the CPU.

assemble_tetra_nedelec_curl_mass_csc_jvp / _vjp
(src/assembly/assembly_tetra_nedelec_3d.f90:645-805, 2114-2287).  The
two-tet fixture, step and bars are the reference's own
(test/test_tetra_nedelec_curl_mass_csc_ad.f90, order 2, a negative mass
coefficient); the other checks hold by the mathematics: under the covariant
map the curl-curl matrix scales like 1/h and the mass matrix like h, so a
dilation x -> (1 + s) x has the tangent -alpha K + beta M; translations and
rotations leave the matrix unchanged (zero tangent, zero net force and
torque of the vertex cotangent); the matrix is linear in alpha and beta.
"""
import numpy as np
import pytest

import oracle as O

FIX_TETS = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], dtype=np.int32)
FIX_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
FIX_VD = np.array([[0.01, -0.02, 0.015], [-0.03, 0.01, 0.02], [0.025, -0.015, 0.005],
                   [-0.01, 0.03, -0.02], [0.02, -0.01, 0.025]])
FIX_A, FIX_B, FIX_AD, FIX_BD = 1.7, -0.8, 0.13, -0.09


def distorted_box(N, seed, amplitude=0.2):
    v, t = O.unit_box_mesh(N)
    rng = np.random.default_rng(seed)
    return v + (amplitude / N) * rng.uniform(-1, 1, v.shape), t


def test_reference_fixture():
    k, step = 2, 2.0e-7
    a = O.nedelec_csc(k, FIX_V, FIX_TETS, FIX_A, FIX_B)
    ad = O.nedelec_csc_jvp(k, FIX_V, FIX_TETS, FIX_A, FIX_B, FIX_VD, FIX_AD, FIX_BD)
    ap = O.nedelec_csc(k, FIX_V + step * FIX_VD, FIX_TETS, FIX_A + step * FIX_AD, FIX_B + step * FIX_BD)
    am = O.nedelec_csc(k, FIX_V - step * FIX_VD, FIX_TETS, FIX_A - step * FIX_AD, FIX_B - step * FIX_BD)
    assert ad.nnz == a.nnz == ap.nnz == am.nnz
    assert np.array_equal(ad.col_ptr, a.col_ptr) and np.array_equal(ad.row_idx, a.row_idx)
    err = np.max(np.abs(ad.val - (ap.val - am.val) / (2 * step))) / max(1.0, np.max(np.abs(ad.val)))
    assert err < 5.0e-8
    e = np.arange(1, a.nnz + 1, dtype=np.float64)
    vb_in = 0.003 * e - 0.00001 * e ** 2
    vb, ab, bb, _ = O.nedelec_csc_vjp(k, FIX_V, FIX_TETS, vb_in, FIX_A, FIX_B)
    lhs = vb_in @ ad.val
    rhs = np.sum(vb * FIX_VD) + ab * FIX_AD + bb * FIX_BD
    assert abs(lhs - rhs) < 3.0e-10 * max(1.0, abs(lhs), abs(rhs))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_rigid_motions_and_dilation(k):
    v, t = distorted_box(2, k)
    A = O.nedelec_csc(k, v, t, 1.3, 0.6)
    Kc = O.nedelec_csc(k, v, t, 1.0, 0.0)
    Mm = O.nedelec_csc(k, v, t, 0.0, 1.0)
    scale = np.max(np.abs(A.val))
    tr = np.tile([0.3, -0.7, 1.1], (v.shape[0], 1))
    assert np.max(np.abs(O.nedelec_csc_jvp(k, v, t, 1.3, 0.6, tr).val)) <= 1e-12 * scale
    rot = np.cross(np.array([0.2, -0.5, 0.9]), v)
    assert np.max(np.abs(O.nedelec_csc_jvp(k, v, t, 1.3, 0.6, rot).val)) <= 1e-12 * scale
    dil = O.nedelec_csc_jvp(k, v, t, 1.3, 0.6, v)
    assert np.max(np.abs(dil.val - (-1.3 * Kc.val + 0.6 * Mm.val))) <= 1e-12 * scale


@pytest.mark.parametrize("k", [1, 2])
def test_vjp_invariants_and_coefficient_bars(k):
    v, t = distorted_box(2, 10 + k)
    Kc = O.nedelec_csc(k, v, t, 1.0, 0.0)
    Mm = O.nedelec_csc(k, v, t, 0.0, 1.0)
    rng = np.random.default_rng(k)
    vb_in = rng.uniform(-1, 1, Kc.nnz)
    vb, ab, bb, cab = O.nedelec_csc_vjp(k, v, t, vb_in, 1.3, 0.6)
    ref = np.abs(vb).sum()
    assert np.max(np.abs(vb.sum(axis=0))) <= 1e-12 * ref
    assert np.max(np.abs(np.cross(v, vb).sum(axis=0))) <= 1e-12 * ref
    assert abs(ab - vb_in @ Kc.val) <= 1e-12 * np.abs(vb_in * Kc.val).sum()
    assert abs(bb - vb_in @ Mm.val) <= 1e-12 * np.abs(vb_in * Mm.val).sum()
    assert abs(cab.sum() - ab) <= 1e-12 * max(1.0, abs(ab))
    # per-cell alpha: tangent in alpha alone is the curl matrix with alpha_dot
    ca = np.exp(rng.normal(size=t.shape[0]))
    cad = rng.normal(size=t.shape[0])
    jd = O.nedelec_csc_jvp(k, v, t, 1.0, 0.6, None, 0.0, 0.0, cell_alpha=ca, cell_alpha_dot=cad)
    _, _, _, cab2 = O.nedelec_csc_vjp(k, v, t, vb_in, 1.0, 0.6, cell_alpha=ca)
    assert abs(cab2 @ cad - vb_in @ jd.val) <= 1e-12 * max(1.0, np.abs(vb_in * jd.val).sum())
