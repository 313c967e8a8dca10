"""The symbolic phase on the device: first-appearance dof maps (radix sort of
keyed occurrences), the CSC pattern (radix sort of element-entry keys) and
the scatter slots, against the oracle on meshes whose vertex labels, cell
order and local vertex order are scrambled (so first appearance differs from
any sorted order).

Checks: global dofs equal to the oracle's build_tetra_lagrange_dof_map /
build_tetra_nedelec_dof_map (tetra_lagrange_global_dof_map.f90:13-79,
tetra_nedelec_global_dof_map.f90:15-105), bit-exact; assembled CSC pattern
bit-exact and values within 1e-12 of the oracle's (the slots put every
element entry in its place).
"""
import numpy as np
import pytest

import oracle as O
from colouring import assert_valid_colouring

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def scrambled(n, seed):
    """unit box mesh with relabelled vertices, shuffled cells and an even
    permutation of every cell's vertices (orientation kept)."""
    v, t = O.unit_box_mesh(n)
    rng = np.random.default_rng(seed)
    lab = rng.permutation(v.shape[0])  # old index -> new index
    v2 = np.empty_like(v)
    v2[lab] = v
    t2 = lab[t - 1] + 1
    t2 = t2[rng.permutation(t2.shape[0])]
    even = np.array([[0, 1, 2, 3], [1, 0, 3, 2], [2, 3, 0, 1], [3, 2, 1, 0],
                     [0, 2, 3, 1], [0, 3, 1, 2], [1, 2, 0, 3], [2, 0, 1, 3]])
    pick = even[rng.integers(0, len(even), t2.shape[0])]
    t2 = np.take_along_axis(t2, pick, axis=1).astype(np.int32)
    return v2, np.ascontiguousarray(t2)


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_lagrange_dofs_on_scrambled_mesh(ff, degree):
    v, t = scrambled(4, degree)
    p = ff.Problem(v, t)
    n, _ = p.pattern_build(ff.capi.FAMILY_H1, degree)
    st, g, G = O.lagrange_dof_map(degree, t)
    assert st == 0 and n == G
    assert np.array_equal(p.global_dofs(base=1), g)
    assert_valid_colouring(p)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_nedelec_dofs_on_scrambled_mesh(ff, k):
    v, t = scrambled(4, 10 + k)
    p = ff.Problem(v, t)
    n, _ = p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
    gdo, eo, fp, nd, ne, nf = O.nedelec_dof_map(k, t)
    assert n == nd and np.array_equal(p.global_dofs(), gdo)
    assert_valid_colouring(p)


@pytest.mark.parametrize("degree", [1, 2])
def test_lagrange_csc_on_scrambled_mesh(ff, degree):
    v, t = scrambled(6, 20 + degree)
    a, st = ff.assemble_tetra_lagrange_stiffness_csc(v, t, degree, 2 * degree, 1.3, 0.4)
    assert st.code == 0
    o, _, _ = O.lagrange_stiffness_csc(v, t, degree, 2 * degree, kappa=1.3, mu=0.4)
    assert np.array_equal(a.col_ptr, o.col_ptr) and np.array_equal(a.row_idx, o.row_idx)
    assert np.max(np.abs(a.val - o.val)) <= 1e-12 * np.max(np.abs(o.val))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_nedelec_csc_on_scrambled_mesh(ff, k):
    v, t = scrambled(4, 30 + k)
    a, st = ff.assemble_tetra_nedelec_curl_mass_csc(v, t, curl_coefficient=0.9, mass_coefficient=1.4,
                                                    order=k)
    assert st.code == 0
    o = O.nedelec_csc(k, v, t, alpha=0.9, beta=1.4)
    assert np.array_equal(a.col_ptr, o.col_ptr) and np.array_equal(a.row_idx, o.row_idx)
    assert np.max(np.abs(a.val - o.val)) <= 1e-12 * np.max(np.abs(o.val))


def _neighbour_degree(inc):
    """Largest number of distinct other cells sharing an entity with one cell
    (incidence rows of entity ids)."""
    T = inc.shape[0]
    ent = inc.reshape(-1)
    cell = np.repeat(np.arange(T), inc.shape[1])
    order = np.argsort(ent, kind="stable")
    ent_s, cell_s = ent[order], cell[order]
    starts = np.flatnonzero(np.r_[True, ent_s[1:] != ent_s[:-1]])
    ends = np.r_[starts[1:], ent_s.shape[0]]
    nb = [set() for _ in range(T)]
    for a, b in zip(starts, ends):
        cs = cell_s[a:b]
        for c in cs:
            nb[c].update(cs)
    return max(len(s) - 1 for s in nb)


@pytest.mark.parametrize("family,order", [(1, 1), (1, 3), (2, 1), (2, 3)])
def test_device_colouring_properties(ff, family, order):
    """The device colouring (Jones-Plassmann rounds, first-fit colours, forced
    by FFH_COLOUR=device on this small mesh) on a scrambled mesh: a proper
    colouring of the dof-sharing graph (the validator), cells ascending
    inside each colour, the same colouring from a second pattern build, and
    no more colours than the first-fit bound of
    one plus the largest number of cells a cell conflicts with (vertex
    sharing for H1, edge sharing for Nedelec).  Nedelec cells are further
    grouped by orientation class inside a colour (ascending inside each
    class), so a colour's list falls at most 23 times there."""
    import os

    v, t = scrambled(5, 40 + 10 * family + order)
    os.environ["FFH_COLOUR"] = "device"  # this mesh is below the device-colouring size
    try:
        p = ff.Problem(v, t)
        p.pattern_build(family, order)
        q = ff.Problem(v, t)
        q.pattern_build(family, order)
    finally:
        del os.environ["FFH_COLOUR"]
    nc = assert_valid_colouring(p)
    cs, cells = p.colouring(base=0)
    for c in range(nc):
        falls = int(np.sum(np.diff(cells[cs[c]:cs[c + 1]]) < 0))
        assert falls == 0 if family == 1 else falls <= 23
    cs2, cells2 = q.colouring(base=0)
    assert np.array_equal(cs, cs2) and np.array_equal(cells, cells2)
    if family == 1:
        inc = t - 1
    else:
        gd = p.global_dofs(base=0)
        inc = gd[:, 0:6 * order:order] // order  # edge numbers (edge dofs lead)
    assert nc <= _neighbour_degree(inc) + 1
