"""Nedelec curl-mass assembly tangents on the GPU against the oracle.

ffh_assemble_jvp / ffh_assemble_vjp on FFH_FAMILY_NEDELEC
(assemble_tetra_nedelec_curl_mass_csc_jvp / _vjp,
src/assembly/assembly_tetra_nedelec_3d.f90:2114-2287).  Bars: 1e-12
relative to the largest entry for the tangent values and the vertex
cotangents (class tables vs the oracle's quadrature loop), 1e-12 for the
coefficient cotangents; the reference test's two-tet fixture with its step
and bars; at a larger size the invariants the oracle tests pin (zero
tangent under translation, -alpha K + beta M under dilation, zero net force
of the vertex cotangent).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FIX_TETS = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], dtype=np.int32)
FIX_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
FIX_VD = np.array([[0.01, -0.02, 0.015], [-0.03, 0.01, 0.02], [0.025, -0.015, 0.005],
                   [-0.01, 0.03, -0.02], [0.02, -0.01, 0.025]])
FIX_A, FIX_B, FIX_AD, FIX_BD = 1.7, -0.8, 0.13, -0.09


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def distorted_box(N, seed, amplitude=0.2):
    v, t = O.unit_box_mesh(N)
    rng = np.random.default_rng(seed)
    return v + (amplitude / N) * rng.uniform(-1, 1, v.shape), t


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_reference_fixture(ff):
    k, step = 2, 2.0e-7
    ad, st = ff.assemble_tetra_nedelec_curl_mass_csc_jvp(FIX_V, FIX_TETS, k, FIX_A, FIX_B, FIX_VD, FIX_AD,
                                                         FIX_BD)
    assert st.code == 0
    ap, _ = ff.assemble_tetra_nedelec_curl_mass_csc(FIX_V + step * FIX_VD, FIX_TETS, FIX_A + step * FIX_AD,
                                                    FIX_B + step * FIX_BD, order=k)
    am, _ = ff.assemble_tetra_nedelec_curl_mass_csc(FIX_V - step * FIX_VD, FIX_TETS, FIX_A - step * FIX_AD,
                                                    FIX_B - step * FIX_BD, order=k)
    assert np.array_equal(ad.col_ptr, ap.col_ptr) and np.array_equal(ad.row_idx, ap.row_idx)
    err = np.max(np.abs(ad.val - (ap.val - am.val) / (2 * step))) / max(1.0, np.max(np.abs(ad.val)))
    assert err < 5.0e-8
    e = np.arange(1, ad.nnz + 1, dtype=np.float64)
    vb_in = 0.003 * e - 0.00001 * e ** 2
    vb, ab, bb, st = ff.assemble_tetra_nedelec_curl_mass_csc_vjp(FIX_V, FIX_TETS, k, FIX_A, FIX_B, vb_in)
    assert st.code == 0
    lhs = vb_in @ ad.val
    rhs = np.sum(vb * FIX_VD) + ab * FIX_AD + bb * FIX_BD
    assert abs(lhs - rhs) < 3.0e-10 * max(1.0, abs(lhs), abs(rhs))
    oref = O.nedelec_csc_jvp(k, FIX_V, FIX_TETS, FIX_A, FIX_B, FIX_VD, FIX_AD, FIX_BD)
    assert relmax(ad.val, oref.val) <= 1e-12


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_jvp_vjp_match_oracle(ff, k):
    v, t = distorted_box(2, k)
    rng = np.random.default_rng(200 + k)
    vd = rng.uniform(-1, 1, v.shape)
    a, b, ad, bd = 1.3, 0.6, 0.2, -0.4
    oj = O.nedelec_csc_jvp(k, v, t, a, b, vd, ad, bd)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    m = p.matrix(ff.capi.FULL).download()
    assert np.array_equal(m.col_ptr, oj.col_ptr) and np.array_equal(m.row_idx, oj.row_idx)
    gj = p.assemble_jvp(-1, kappa=a, mass=b, vertices_dot=vd, kappa_dot=ad, mass_dot=bd)
    assert relmax(gj, oj.val) <= 1e-12
    vb_in = rng.uniform(-1, 1, oj.nnz)
    ovb, oab, obb, _ = O.nedelec_csc_vjp(k, v, t, vb_in, a, b)
    gvb, gab, gbb = p.assemble_vjp(-1, vb_in, kappa=a, mass=b)
    assert relmax(gvb, ovb) <= 1e-12
    # vertex cotangents are summed per vertex in cell order: bit-for-bit repeatable
    gvb2, _, _ = p.assemble_vjp(-1, vb_in, kappa=a, mass=b)
    assert np.array_equal(gvb, gvb2)
    assert abs(gab - oab) <= 1e-12 * max(1.0, abs(oab))
    assert abs(gbb - obb) <= 1e-12 * max(1.0, abs(obb))


def test_cell_alpha_tangents_match_oracle(ff):
    v, t = distorted_box(2, 9)
    k = 2
    rng = np.random.default_rng(10)
    ca = np.exp(rng.normal(size=t.shape[0]))
    cad = rng.normal(size=t.shape[0])
    vd = rng.uniform(-1, 1, v.shape)
    oj = O.nedelec_csc_jvp(k, v, t, 1.0, 0.3, vd, 0.0, 0.1, cell_alpha=ca, cell_alpha_dot=cad)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    gj = p.assemble_jvp(-1, cell_kappa=ca, mass=0.3, vertices_dot=vd, cell_kappa_dot=cad, mass_dot=0.1)
    assert relmax(gj, oj.val) <= 1e-12
    vb_in = rng.uniform(-1, 1, oj.nnz)
    ovb, _, obb, ocab = O.nedelec_csc_vjp(k, v, t, vb_in, 1.0, 0.3, cell_alpha=ca)
    gvb, gcab, gbb = p.assemble_vjp(-1, vb_in, cell_kappa=ca, mass=0.3)
    assert relmax(gvb, ovb) <= 1e-12
    assert relmax(gcab, ocab) <= 1e-12
    assert abs(gbb - obb) <= 1e-12 * max(1.0, abs(obb))


def test_invariants_at_size(ff):
    """k = 2 on a distorted 16^3 box (24576 tets)."""
    v, t = distorted_box(16, 3, amplitude=0.1)
    k = 2
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    p.assemble(-1, kappa=1.0, mass=0.0)
    Kc = p.matrix(ff.capi.FULL).download().val.copy()
    p.assemble(-1, kappa=0.0, mass=1.0)
    Mm = p.matrix(ff.capi.FULL).download().val.copy()
    scale = max(np.max(np.abs(Kc)), np.max(np.abs(Mm)))
    tr = np.tile([0.3, -0.7, 1.1], (v.shape[0], 1))
    assert np.max(np.abs(p.assemble_jvp(-1, kappa=1.3, mass=0.6, vertices_dot=tr))) <= 1e-11 * scale
    dil = p.assemble_jvp(-1, kappa=1.3, mass=0.6, vertices_dot=v)
    assert np.max(np.abs(dil - (-1.3 * Kc + 0.6 * Mm))) <= 1e-11 * scale
    vb_in = np.random.default_rng(4).uniform(-1, 1, Kc.shape[0])
    vb, ab, bb = p.assemble_vjp(-1, vb_in, kappa=1.3, mass=0.6)
    assert np.max(np.abs(vb.sum(axis=0))) <= 1e-10 * np.abs(vb).sum()
    assert abs(ab - vb_in @ Kc) <= 1e-11 * np.abs(vb_in * Kc).sum()
    assert abs(bb - vb_in @ Mm) <= 1e-11 * np.abs(vb_in * Mm).sum()
