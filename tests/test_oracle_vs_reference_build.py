"""The oracle against reference modules compiled from their own sources
(oracle/_ref, built by oracle/ref.mk in this container only)."""
import numpy as np
import pytest

import oracle as O
from oracle import ref as R

pytestmark = pytest.mark.skipif(not R.available(), reason="reference build (oracle/_ref) not present")


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_lagrange_basis_equals_reference(degree):
    rng = np.random.default_rng(degree)
    for _ in range(20):
        p = rng.dirichlet(np.ones(4))[1:]
        s1, v1, g1 = O.lagrange_eval(degree, p)
        s2, v2, g2 = R.lagrange_eval(degree, p)
        assert s1 == s2 == 0
        assert np.max(np.abs(v1 - v2)) < 1e-14
        assert np.max(np.abs(g1 - g2)) < 1e-13


@pytest.mark.parametrize("degree,N", [(1, 3), (2, 3), (3, 2), (4, 2)])
def test_lagrange_dof_map_equals_reference(degree, N):
    v, t = O.unit_box_mesh(N)
    s1, g1, c1 = O.lagrange_dof_map(degree, t)
    s2, g2, c2 = R.lagrange_dof_map(degree, t)
    assert s1 == s2 == 0 and c1 == c2
    assert np.array_equal(g1, g2)


def test_lagrange_dof_map_rejects_like_reference():
    bad = np.array([[1, 2, 2, 3]], dtype=np.int32)
    assert R.lagrange_dof_map(1, bad)[0] != 0 and O.lagrange_dof_map(1, bad)[0] != 0


def test_generated_h1_oracle_kat():
    """test/test_tetra_h1_manufactured_oracle.f90:12-23"""
    v = R.h1_oracle(0.1, 0.2, 0.3)
    assert abs(v[0] - 0.0024) < 1e-15
    assert np.max(np.abs(v[1:4] - [0.018, 0.006, 0.002])) < 1e-15
    assert abs(v[4] - 0.22) < 1e-14
