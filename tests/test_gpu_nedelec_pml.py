"""Complex PML Nedelec assembly on the GPU (ffh_nedelec_pml_values) against
the oracle: pattern bit-exact, real and imaginary values to 1e-12 of the
largest magnitude, cartesian (assembly_tetra_nedelec_3d.f90:1613-1685) and
curvilinear (:1372-1444) stretches, orders 1-4; the reference's
constant-field energy (test/test_tetra_nedelec_pml_sparse.f90:20-51) through
the product path; invalid stretches rejected with a status.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_nedelec_pml import REF_T, REF_V, dense, masses, whitney_dofs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def mesh(n, seed):
    v, t = O.unit_box_mesh(n)
    return v + (0.1 / n) * np.random.default_rng(seed).uniform(-1, 1, v.shape), t


def same(a, o):
    assert np.array_equal(a.col_ptr, o.col_ptr) and np.array_equal(a.row_idx, o.row_idx)
    scale = np.max(np.abs(o.val))
    assert np.max(np.abs(a.val.real - o.val.real)) <= 1e-12 * scale
    assert np.max(np.abs(a.val.imag - o.val.imag)) <= 1e-12 * scale


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_cartesian_matches_oracle(ff, k):
    v, t = mesh(2, k)
    rng = np.random.default_rng(10 + k)
    s = rng.uniform(0.6, 1.4, (t.shape[0], 3)) + 1j * rng.uniform(-0.6, 0.6, (t.shape[0], 3))
    a, st = ff.assemble_tetra_nedelec_pml_csc(v, t, k, s, 1.7)
    assert st.code == 0
    same(a, O.nedelec_pml_csc(k, v, t, s, 1.7))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_curvilinear_matches_oracle(ff, k):
    v, t = mesh(2, 20 + k)
    rng = np.random.default_rng(30 + k)
    S = (np.eye(3) + rng.uniform(-0.3, 0.3, (t.shape[0], 3, 3))
         + 1j * rng.uniform(-0.4, 0.4, (t.shape[0], 3, 3)))
    a, st = ff.assemble_tetra_nedelec_curvilinear_pml_csc(v, t, k, S, 0.8)
    assert st.code == 0
    same(a, O.nedelec_pml_csc(k, v, t, S, 0.8, kind=1))


def test_larger_mesh_matches_oracle(ff):
    v, t = mesh(6, 5)
    rng = np.random.default_rng(6)
    s = rng.uniform(0.6, 1.4, (t.shape[0], 3)) + 1j * rng.uniform(-0.6, 0.6, (t.shape[0], 3))
    a, st = ff.assemble_tetra_nedelec_pml_csc(v, t, 2, s, 2.5)
    assert st.code == 0
    same(a, O.nedelec_pml_csc(2, v, t, s, 2.5))


def test_reference_constant_field_energy(ff):
    s = np.array([[1 + 0.6j, 1, 1]] * 2)
    a, st = ff.assemble_tetra_nedelec_pml_csc(REF_V, REF_T, 1, s, 1.3)
    assert st.code == 0
    f = np.array([1.0, 2.0, -1.0])
    d = whitney_dofs(REF_V, REF_T, lambda x: f)
    energy = d @ (dense(a) @ d)
    assert abs(energy - (-1.3 ** 2 / 3.0 * np.sum(masses(s[0]) * f ** 2))) < 3.0e-12
    for k in (2, 3, 4):
        a, st = ff.assemble_tetra_nedelec_pml_csc(REF_V, REF_T, k, s, 1.3)
        assert st.code == 0 and a.nrow > 0


def test_rejects_invalid_stretch(ff):
    s = np.array([[1.0, 0.0, 1.0]] * 2, dtype=complex)
    a, st = ff.assemble_tetra_nedelec_pml_csc(REF_V, REF_T, 1, s, 1.3)
    assert a is None and st.code != 0
    S = np.zeros((2, 3, 3), dtype=complex)
    S[:, 0, 0] = S[:, 1, 1] = 1.0
    a, st = ff.assemble_tetra_nedelec_curvilinear_pml_csc(REF_V, REF_T, 1, S, 1.3)
    assert a is None and st.code != 0
    a, st = ff.assemble_tetra_nedelec_pml_csc(REF_V, REF_T, 1, np.ones((2, 3)), -1.0)
    assert a is None and st.code != 0


def _fd(k, v, t, S, k0, vd, sd, kd, kind, h=1e-6):
    ap = O.nedelec_pml_csc(k, v + h * vd, t, S + h * sd, k0 + h * kd, kind=kind)
    am = O.nedelec_pml_csc(k, v - h * vd, t, S - h * sd, k0 - h * kd, kind=kind)
    return ap, (ap.val - am.val) / (2 * h)


@pytest.mark.parametrize("k", [1, 2])
def test_cartesian_jvp_matches_central_differences(ff, k):
    """assemble_tetra_nedelec_pml_csc_jvp (:1687-1768) against central
    differences of the oracle's primal along (vertices, stretch, wave number)."""
    v, t = mesh(2, 40 + k)
    rng = np.random.default_rng(50 + k)
    s = rng.uniform(0.6, 1.4, (t.shape[0], 3)) + 1j * rng.uniform(-0.6, 0.6, (t.shape[0], 3))
    vd = rng.uniform(-1, 1, v.shape)
    sd = rng.uniform(-1, 1, s.shape) + 1j * rng.uniform(-1, 1, s.shape)
    a, st = ff.assemble_tetra_nedelec_pml_csc_jvp(v, t, k, s, 1.7, vd, sd, 0.3)
    assert st.code == 0
    o, fd = _fd(k, v, t, s, 1.7, vd, sd, 0.3, 0)
    assert np.array_equal(a.col_ptr, o.col_ptr) and np.array_equal(a.row_idx, o.row_idx)
    assert np.max(np.abs(a.val - fd)) <= 2e-7 * np.max(np.abs(fd))
    # each direction alone, and linearity in the tangent
    parts = [ff.assemble_tetra_nedelec_pml_csc_jvp(v, t, k, s, 1.7, vd if i == 0 else np.zeros_like(vd),
                                                   sd if i == 1 else np.zeros_like(sd), 0.3 if i == 2 else 0.0)[0].val
             for i in range(3)]
    assert np.max(np.abs(sum(parts) - a.val)) <= 1e-12 * np.max(np.abs(a.val))


@pytest.mark.parametrize("k", [1, 2])
def test_curvilinear_jvp_matches_central_differences(ff, k):
    """assemble_tetra_nedelec_curvilinear_pml_csc_jvp (:1446-1521)."""
    v, t = mesh(2, 60 + k)
    rng = np.random.default_rng(70 + k)
    S = (np.eye(3) + rng.uniform(-0.3, 0.3, (t.shape[0], 3, 3))
         + 1j * rng.uniform(-0.4, 0.4, (t.shape[0], 3, 3)))
    vd = rng.uniform(-1, 1, v.shape)
    sd = rng.uniform(-1, 1, S.shape) + 1j * rng.uniform(-1, 1, S.shape)
    a, st = ff.assemble_tetra_nedelec_curvilinear_pml_csc_jvp(v, t, k, S, 0.8, vd, sd, -0.2)
    assert st.code == 0
    o, fd = _fd(k, v, t, S, 0.8, vd, sd, -0.2, 1)
    assert np.array_equal(a.row_idx, o.row_idx)
    assert np.max(np.abs(a.val - fd)) <= 2e-7 * np.max(np.abs(fd))


@pytest.mark.parametrize("kind,k", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_vjp_is_the_adjoint_of_the_jvp(ff, kind, k):
    """assemble_tetra_nedelec_(curvilinear_)pml_csc_vjp (:1770-1868,
    :1523-1611) in the reference's convention (cotangent = sum seed *
    conj(d value / d input), real part for real inputs):
    Re sum conj(w) A_dot = <v_bar, v_dot> + Re sum conj(s_bar) s_dot + k_bar k_dot
    for the JVP checked above against central differences."""
    v, t = mesh(2, 80 + k)
    rng = np.random.default_rng(90 + 10 * kind + k)
    shape = (t.shape[0], 3) if kind == 0 else (t.shape[0], 3, 3)
    S = rng.uniform(0.6, 1.4, shape) + 1j * rng.uniform(-0.5, 0.5, shape)
    if kind == 1:
        S = S * 0.3 + np.eye(3)
    vd = rng.uniform(-1, 1, v.shape)
    sd = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    kd = 0.37
    jvp = ff.assemble_tetra_nedelec_pml_csc_jvp if kind == 0 else ff.assemble_tetra_nedelec_curvilinear_pml_csc_jvp
    vjp = ff.assemble_tetra_nedelec_pml_csc_vjp if kind == 0 else ff.assemble_tetra_nedelec_curvilinear_pml_csc_vjp
    a, st = jvp(v, t, k, S, 1.3, vd, sd, kd)
    assert st.code == 0
    w = rng.uniform(-1, 1, a.val.shape) + 1j * rng.uniform(-1, 1, a.val.shape)
    vb, sb, kb, st = vjp(v, t, k, S, 1.3, w)
    assert st.code == 0
    lhs = np.real(np.sum(np.conj(w) * a.val))
    rhs = np.sum(vb * vd) + np.real(np.sum(np.conj(sb) * sd)) + kb * kd
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)
    # each input separately
    for i in range(3):
        ai, _ = jvp(v, t, k, S, 1.3, vd if i == 0 else 0 * vd, sd if i == 1 else 0 * sd, kd if i == 2 else 0.0)
        part = [np.sum(vb * vd), np.real(np.sum(np.conj(sb) * sd)), kb * kd][i]
        li = np.real(np.sum(np.conj(w) * ai.val))
        assert abs(li - part) <= 1e-10 * max(abs(li), 1.0), i


def test_pml_system_solved_by_cocg(ff):
    """The device PML operator solved by ComplexMatrix (Jacobi-COCG): the
    residual of the oracle's matrix at the computed x is below 1e-9 ||b||."""
    v, t = mesh(2, 99)
    rng = np.random.default_rng(99)
    s = rng.uniform(0.8, 1.2, (t.shape[0], 3)) + 1j * rng.uniform(0.2, 0.6, (t.shape[0], 3))
    a, st = ff.assemble_tetra_nedelec_pml_csc(v, t, 2, s, 0.9)
    assert st.code == 0
    b = rng.uniform(-1, 1, a.nrow) + 1j * rng.uniform(-1, 1, a.nrow)
    m = ff.ComplexMatrix(a)
    x, it, rel, conv = m.solve(b, tolerance=1e-11, max_iterations=50000)
    assert conv and rel <= 1e-11
    A = dense(O.nedelec_pml_csc(2, v, t, s, 0.9))
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)
    m.close()
