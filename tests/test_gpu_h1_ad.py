"""H1 assembly tangents on the GPU against the oracle.

ffh_assemble_jvp / ffh_assemble_vjp (assemble_tetra_lagrange_stiffness_csc_jvp
/ _vjp, src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:336-471)
through the C-ABI.  Bars: 1e-12 relative to the largest entry for the
tangent values and the vertex cotangents (reference-tensor form vs the
oracle's quadrature loop), 1e-12 for the coefficient cotangents; the
reference test's own two-tet fixture with its step and bars; and, at a
larger size, the invariants the oracle tests pin (zero tangent under
translation, K + 3M under dilation, zero net force of the vertex
cotangent).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FIX_TETS = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], dtype=np.int32)
FIX_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
FIX_VD = np.array([[0.01, -0.02, 0.015], [-0.03, 0.01, 0.02], [0.025, -0.015, 0.005],
                   [-0.01, 0.03, -0.02], [0.02, -0.01, 0.025]])
FIX_K, FIX_M, FIX_KD, FIX_MD = 1.7, 0.8, 0.13, -0.09


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def distorted_box(N, seed, amplitude=0.25):
    v, t = O.unit_box_mesh(N)
    rng = np.random.default_rng(seed)
    return v + (amplitude / N) * rng.uniform(-1, 1, v.shape), t


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_reference_fixture(ff):
    d, q, step = 2, 6, 2.0e-7
    ad, st = ff.assemble_tetra_lagrange_stiffness_csc_jvp(FIX_V, FIX_TETS, d, q, FIX_K, FIX_M, FIX_VD,
                                                          FIX_KD, FIX_MD)
    assert st.code == 0
    ap, _ = ff.assemble_tetra_lagrange_stiffness_csc(FIX_V + step * FIX_VD, FIX_TETS, d, q,
                                                     FIX_K + step * FIX_KD, FIX_M + step * FIX_MD)
    am, _ = ff.assemble_tetra_lagrange_stiffness_csc(FIX_V - step * FIX_VD, FIX_TETS, d, q,
                                                     FIX_K - step * FIX_KD, FIX_M - step * FIX_MD)
    assert np.array_equal(ad.col_ptr, ap.col_ptr) and np.array_equal(ad.row_idx, ap.row_idx)
    err = np.max(np.abs(ad.val - (ap.val - am.val) / (2 * step))) / max(1.0, np.max(np.abs(ad.val)))
    assert err < 5.0e-8
    e = np.arange(1, ad.nnz + 1, dtype=np.float64)
    vb_in = 0.003 * e - 0.00001 * e ** 2
    vb, kb, mb, st = ff.assemble_tetra_lagrange_stiffness_csc_vjp(FIX_V, FIX_TETS, d, q, FIX_K, FIX_M, vb_in)
    assert st.code == 0
    lhs = vb_in @ ad.val
    rhs = np.sum(vb * FIX_VD) + kb * FIX_KD + mb * FIX_MD
    assert abs(lhs - rhs) < 3.0e-10 * max(1.0, abs(lhs), abs(rhs))
    oref = O.lagrange_stiffness_csc_jvp(FIX_V, FIX_TETS, d, q, FIX_K, FIX_M, FIX_VD, FIX_KD, FIX_MD)
    assert relmax(ad.val, oref.val) <= 1e-12


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_jvp_vjp_match_oracle(ff, d):
    v, t = distorted_box(3, d)
    q = 2 * d
    rng = np.random.default_rng(100 + d)
    vd = rng.uniform(-1, 1, v.shape)
    k, mu, kd, mud = 1.3, 0.6, 0.2, -0.4
    oj = O.lagrange_stiffness_csc_jvp(v, t, d, q, k, mu, vd, kd, mud)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, d)
    a = p.matrix(ff.capi.FULL).download()
    assert np.array_equal(a.col_ptr, oj.col_ptr) and np.array_equal(a.row_idx, oj.row_idx)
    gj = p.assemble_jvp(q, kappa=k, mass=mu, vertices_dot=vd, kappa_dot=kd, mass_dot=mud)
    assert relmax(gj, oj.val) <= 1e-12
    vb_in = rng.uniform(-1, 1, oj.nnz)
    ovb, okb, omb, _ = O.lagrange_stiffness_csc_vjp(v, t, d, q, vb_in, k, mu)
    gvb, gkb, gmb = p.assemble_vjp(q, vb_in, kappa=k, mass=mu)
    assert relmax(gvb, ovb) <= 1e-12
    assert abs(gkb - okb) <= 1e-12 * max(1.0, abs(okb))
    assert abs(gmb - omb) <= 1e-12 * max(1.0, abs(omb))
    # the primal matrix of the same problem is unchanged by the tangent calls
    p.assemble(q, kappa=k, mass=mu)
    ao, _, _ = O.lagrange_stiffness_csc(v, t, d, q, kappa=k, mu=mu)
    assert relmax(p.matrix(ff.capi.FULL).download().val, ao.val) <= 1e-12


def test_cell_kappa_tangents_match_oracle(ff):
    v, t = distorted_box(3, 7)
    d, q = 2, 4
    rng = np.random.default_rng(8)
    ck = np.exp(rng.normal(size=t.shape[0]))
    ckd = rng.normal(size=t.shape[0])
    vd = rng.uniform(-1, 1, v.shape)
    oj = O.lagrange_stiffness_csc_jvp(v, t, d, q, 1.0, 0.3, vd, 0.0, 0.1, cell_kappa=ck, cell_kappa_dot=ckd)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, d)
    gj = p.assemble_jvp(q, cell_kappa=ck, mass=0.3, vertices_dot=vd, cell_kappa_dot=ckd, mass_dot=0.1)
    assert relmax(gj, oj.val) <= 1e-12
    vb_in = rng.uniform(-1, 1, oj.nnz)
    ovb, _, omb, ockb = O.lagrange_stiffness_csc_vjp(v, t, d, q, vb_in, 1.0, 0.3, cell_kappa=ck)
    gvb, gckb, gmb = p.assemble_vjp(q, vb_in, cell_kappa=ck, mass=0.3)
    assert relmax(gvb, ovb) <= 1e-12
    assert relmax(gckb, ockb) <= 1e-12
    assert abs(gmb - omb) <= 1e-12 * max(1.0, abs(omb))


def test_invariants_at_size(ff):
    """P2 on a distorted 24^3 box (82944 tets): translation gives a zero
    tangent, dilation gives kappa K + 3 mu M, and the vertex cotangent has
    zero net force."""
    v, t = distorted_box(24, 3, amplitude=0.1)
    d, q = 2, 4
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, d)
    p.assemble(q, kappa=1.0, mass=0.0)
    K = p.matrix(ff.capi.FULL).download().val.copy()
    p.assemble(q, kappa=0.0, mass=1.0)
    M = p.matrix(ff.capi.FULL).download().val.copy()
    scale = np.max(np.abs(K))
    tr = np.tile([0.3, -0.7, 1.1], (v.shape[0], 1))
    assert np.max(np.abs(p.assemble_jvp(q, kappa=1.3, mass=0.6, vertices_dot=tr))) <= 1e-11 * scale
    dil = p.assemble_jvp(q, kappa=1.3, mass=0.6, vertices_dot=v)
    assert np.max(np.abs(dil - (1.3 * K + 1.8 * M))) <= 1e-11 * scale
    vb_in = np.random.default_rng(4).uniform(-1, 1, K.shape[0])
    vb, kb, mb = p.assemble_vjp(q, vb_in, kappa=1.3, mass=0.6)
    assert np.max(np.abs(vb.sum(axis=0))) <= 1e-10 * np.abs(vb).sum()
    assert abs(kb - vb_in @ K) <= 1e-11 * np.abs(vb_in * K).sum()
    assert abs(mb - vb_in @ M) <= 1e-11 * np.abs(vb_in * M).sum()


def test_errors(ff):
    v, t = distorted_box(2, 1)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, 1)
    with pytest.raises(ValueError):
        p.assemble_vjp(2, np.zeros(3))
    with pytest.raises(ValueError):
        p.assemble_jvp(2, kappa=1.0, cell_kappa_dot=np.zeros(t.shape[0]))
    n = ff.Problem(v, t)  # no pattern yet
    with pytest.raises(ff.FFHError):
        n.assemble_jvp(2, kappa=1.0)
