"""GPU parity of the tetrahedral Nedelec path (HIP through the C-ABI) against
the oracle.

Tolerances (north star): index arrays bit-exact; FP64 values within 1e-12 of
the largest entry; the FP32 contraction reports its error against FP64
(bounded here by 1e-5 relative Frobenius).  Orders 5 and 6 compare against
committed oracle fixtures (tests/golden/nedelec_single_tet.npz) because the
oracle's order-6 element takes ~20 s on the CPU.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def renumbered_box(N, seed):
    """unit_box_mesh(N) with vertex ids relabelled by a random permutation,
    so cells fall in many orientation classes (local order and geometry
    unchanged)."""
    v, t = O.unit_box_mesh(N)
    perm = np.random.default_rng(seed).permutation(v.shape[0])
    v2 = np.zeros_like(v)
    v2[perm] = v
    return v2, (perm[t - 1] + 1).astype(np.int32)


def spd_tensors(T, seed):
    """Q diag(exp(0.5 z)) Q^T with a uniformly random rotation per cell."""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(T, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    Q = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    d = np.exp(0.5 * rng.normal(size=(T, 3)))
    return np.einsum("tij,tj,tkj->tik", Q, d, Q)


def check_csc(a, ref, tol=1e-12):
    assert a.nrow == ref.n_rows and a.nnz == ref.nnz
    assert np.array_equal(a.col_ptr, ref.col_ptr)
    assert np.array_equal(a.row_idx, ref.row_idx)
    err = np.max(np.abs(a.val - ref.val)) / np.max(np.abs(ref.val))
    assert err <= tol, err
    return err


def test_dof_map_matches_oracle(ff):
    v, t = renumbered_box(2, 1)
    for k in range(1, 7):
        p = ff.Problem(v, t)
        n, nnz = p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
        gd = p.global_dofs()
        gdo, eo, fp, nd, ne, nf = O.nedelec_dof_map(k, t)
        assert n == nd and np.array_equal(gd, gdo)
        from colouring import assert_valid_colouring

        assert_valid_colouring(p)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_curl_mass_csc_parity(ff, k):
    v, t = renumbered_box(2 if k <= 3 else 1, k)
    a, st = ff.assemble_tetra_nedelec_curl_mass_csc(v, t, 1.7, 0.8, order=k)
    assert st.code == 0
    check_csc(a, O.nedelec_csc(k, v, t, alpha=1.7, beta=0.8))


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_weighted_csc_parity(ff, k):
    v, t = renumbered_box(1, 10 + k)
    K = spd_tensors(t.shape[0], k)
    a, st = ff.assemble_tetra_nedelec_weighted_csc(v, t, K, 0.6, order=k)
    assert st.code == 0
    check_csc(a, O.nedelec_csc(k, v, t, beta=0.6, cell_tensor=K))


def test_per_cell_alpha_parity(ff):
    """Per-cell curl coefficient alpha_e equals the weighted element with
    K_e = alpha_e I (both rules integrate the degree-2k products exactly)."""
    v, t = renumbered_box(2, 7)
    alpha = np.exp(np.random.default_rng(3).normal(size=t.shape[0]))
    a, st = ff.assemble_tetra_nedelec_curl_mass_csc(v, t, None, 0.9, order=2, cell_coefficient=alpha)
    assert st.code == 0
    K = alpha[:, None, None] * np.eye(3)[None]
    check_csc(a, O.nedelec_csc(2, v, t, beta=0.9, cell_tensor=K))


@pytest.mark.parametrize("name,k,weighted", [("k5", 5, False), ("k6", 6, False), ("k5w", 5, True)])
def test_high_order_single_tet_golden(ff, name, k, weighted):
    import sys

    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_nedelec_golden import TENSOR, single_tet

    g = np.load(os.path.join(HERE, "golden", "nedelec_single_tet.npz"))
    v, t = single_tet()
    if weighted:
        a, st = ff.assemble_tetra_nedelec_weighted_csc(v, t, TENSOR, 0.7, order=k)
    else:
        a, st = ff.assemble_tetra_nedelec_curl_mass_csc(v, t, 1.3, 0.7, order=k)
    assert st.code == 0
    ref = O.CSC(a.nrow, a.ncol, g[name + "_col_ptr"], g[name + "_row_idx"], g[name + "_val"])
    check_csc(a, ref)


@pytest.mark.parametrize("name", ["two_k5", "two_k6", "two_k5w", "two_k6w", "box_k4w", "box_k6w"])
def test_high_order_fixtures(ff, name):
    """Orders 4-6, entry by entry, against the oracle's CSCs committed in
    tests/golden/nedelec_high_order.npz (make_nedelec_high_order.py):

    * the reference's two-tet mesh (test/test_tetra_nedelec_sparse_assembly_slow.f90:
      31-85), whose cells are in two orientation classes and share the face
      1-2-3, so the multi-cell scatter and the shared face's orientation
      transform are checked at k = 5 and 6, curl-mass and tensor-weighted;
    * the relabelled 1^3 x 6 box with seeded per-cell SPD tensors at k = 4 and
      k = 6, the tensor-weighted workload's own orders.

    At k = 6 the oracle's face and interior moments use this build's
    Dubiner / Koornwinder normalisation; the reference's (fortnum's) is not
    available, so k = 6 entries are parity-unpinned against the reference
    itself (the basis-independent checks of test_k6_basis_spans_the_space
    and the constant-field energy hold)."""
    g = np.load(os.path.join(HERE, "golden", "nedelec_high_order.npz"))
    mesh = "two" if name.startswith("two") else "box"
    v, t = g[mesh + "_v"], g[mesh + "_t"]
    k = int(name.split("_k")[1][0])
    if name.endswith("w"):
        beta = 0.7 if mesh == "two" else 0.6
        a, st = ff.assemble_tetra_nedelec_weighted_csc(v, t, g[mesh + "_K"], beta, order=k)
    else:
        a, st = ff.assemble_tetra_nedelec_curl_mass_csc(v, t, 1.3, 0.7, order=k)
    assert st.code == 0
    ref = O.CSC(a.nrow, a.ncol, g[name + "_col_ptr"], g[name + "_row_idx"], g[name + "_val"])
    check_csc(a, ref)


@pytest.mark.parametrize("k", [5, 6])
def test_constant_field_energy(ff, k):
    """test_tetra_nedelec_sparse_assembly_slow.f90:62-85: the interpolant of a
    constant field has energy beta |c|^2 |Omega| = 3 on the two-tet mesh
    (alpha = 4 does not enter: the field is curl-free)."""
    from test_oracle_nedelec import MESH_T, MESH_V, _constant_field_dofs

    a, st = ff.assemble_tetra_nedelec_curl_mass_csc(MESH_V, MESH_T, 4.0, 1.5, order=k)
    dofs = _constant_field_dofs(k, MESH_V, MESH_T, np.array([1.0, 2.0, -1.0]))
    m = ff.Matrix.upload(a)
    assert abs(dofs @ m.spmv(dofs) - 3.0) < 1e-8


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_fp32_weighted_contraction_error(ff, k):
    """The tensor-weighted workload's FP32 element contraction against its
    FP64 one, per order (per-cell SPD tensors, beta = 1): the relative
    Frobenius and max-entry errors are reported by the benchmark and bounded
    here by 1e-6."""
    v, t = renumbered_box(2 if k <= 3 else 1, 20 + k)
    K = spd_tensors(t.shape[0], 30 + k)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    p.assemble(-1, mass=1.0, cell_tensor=K, precision="fp64")
    a64 = p.matrix(ff.capi.FULL).download()
    p.assemble(-1, mass=1.0, reuse_tensor=True, precision="fp32")
    a32 = p.matrix(ff.capi.FULL).download()
    assert np.array_equal(a32.row_idx, a64.row_idx)
    fro = np.linalg.norm(a32.val - a64.val) / np.linalg.norm(a64.val)
    mx = np.max(np.abs(a32.val - a64.val)) / np.max(np.abs(a64.val))
    assert 0 < fro < 1e-6 and mx < 1e-6, (fro, mx)


@pytest.mark.parametrize("k", [1, 3])
def test_fp32_contraction_error(ff, k):
    """The FP32 element contraction and scatter against the FP64 path."""
    v, t = renumbered_box(3, 5)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    p.assemble(-1, kappa=1.3, mass=0.7, precision="fp64")
    a64 = p.matrix(ff.capi.FULL).download()
    p.assemble(-1, kappa=1.3, mass=0.7, precision="fp32")
    a32 = p.matrix(ff.capi.FULL).download()
    assert np.array_equal(a32.row_idx, a64.row_idx)
    fro = np.linalg.norm(a32.val - a64.val) / np.linalg.norm(a64.val)
    assert 0 < fro < 1e-5, fro


def test_zero_tangential_solve_parity(ff):
    """Curl-mass system with zero tangential trace (build_zero_tangential_mask,
    fortfem_tetra_nedelec_solver_3d.f90:431-470) solved by Jacobi-PCG against
    the oracle's reduction and PCG.

    This system takes ~800 iterations; over that many steps CG's iteration
    count moves with the summation order of the dot products (the GPU sums
    per-block partials, the oracle left to right), so the count is held to
    2 % of the oracle's and the solution to 1e-9."""
    v, t = renumbered_box(3, 9)
    k = 2
    p = ff.Problem(v, t)
    G, _ = p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    p.assemble(-1, kappa=1.0, mass=1.0)
    rhs = np.random.default_rng(4).uniform(-1, 1, G)
    p.load(0, ff.capi.SOURCE_DOF_VECTOR, rhs)
    p.dirichlet(0)
    cons = p.constrained()
    a_ref = O.nedelec_csc(k, v, t, alpha=1.0, beta=1.0)
    cons_ref = _tangential_mask(k, v, t)
    assert np.array_equal(cons, cons_ref)
    red, free, rb = O.constrained_reduction(a_ref, cons_ref.astype(np.int32), np.zeros(G), rhs)
    opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    x, st = p.solve(opts)
    xo, sto = O.solve_sparse(red, rb, method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    assert st.converged and abs(st.iterations - sto.iterations) <= max(1, 0.02 * sto.iterations)
    xr = np.zeros(G)
    xr[free - 1] = xo
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)


def _tangential_mask(k, v, t):
    """Literal restatement of build_zero_tangential_mask on a small mesh."""
    gd, eo, fp, nd, ne, nf = O.nedelec_dof_map(k, t)
    st, edges, egd, eor = O.edge_dof_map(t)
    faces = {}
    for c in range(t.shape[0]):
        for f in [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]:
            key = tuple(sorted(t[c, list(f)]))
            if key not in faces:
                faces[key] = len(faces)
    mask = np.zeros(nd, dtype=bool)
    fdc = k * (k - 1)
    for key, fi in faces.items():
        inc = sum(all(x in t[c] for x in key) for c in range(t.shape[0]))
        if inc != 1:
            continue
        for e, (a, b) in enumerate(edges):
            if a in key and b in key:
                mask[e * k:(e + 1) * k] = True
        mask[ne * k + fi * fdc: ne * k + (fi + 1) * fdc] = True
    return mask


@pytest.mark.parametrize("k", [1, 2, 3])
def test_vector_load_parity(ff, k):
    """assemble_tetra_nedelec_vector_load(_order) on the GPU (sampled source at
    the reference's load points, oriented class tables, colour scatter)
    against the oracle, on a mesh with every orientation class present."""
    v, t = renumbered_box(2 if k <= 2 else 1, k)
    f = lambda x, y, z: (np.sin(x) + y, y * z - 0.5, np.cos(z) * x)
    b, st = ff.assemble_tetra_nedelec_vector_load(v, t, f, order=k)
    assert st.code == 0
    bo = O.nedelec_vector_load(k, v, t, lambda x, y, z: np.array(f(x, y, z)))
    assert b.shape == bo.shape
    assert np.max(np.abs(b - bo)) <= 1e-12 * np.max(np.abs(bo))
    # a constant vector through FFH_SOURCE_CONSTANT gives the same as samples
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    p.load(-1, ff.capi.SOURCE_CONSTANT, [0.3, -1.1, 2.0])
    bc = p.rhs(ff.capi.FULL)
    bco = O.nedelec_vector_load(k, v, t, lambda x, y, z: np.array([0.3, -1.1, 2.0]))
    assert np.max(np.abs(bc - bco)) <= 1e-12 * np.max(np.abs(bco))
