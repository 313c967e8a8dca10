"""Pins the oracle's Nedelec vector load (assemble_tetra_nedelec_vector_load,
src/assembly/assembly_tetra_nedelec_3d.f90:2002-2040, and _order, :71-127).

A constant field lies in every first-kind Nedelec space, so its load vector b
satisfies b = M d for the dofs d of the field, with M the mass matrix (the
curl-mass assembly with zero curl coefficient, pinned by the reference's
energy tests): solving M d = b and evaluating the field of d anywhere in the
mesh must give the constant back.  Linearity in the source and the
zero-source case complete the check; a dropped determinant, weight, Piola
transpose or orientation fails the first test.
"""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("k", [1, 2, 3])
def test_constant_field_load_is_mass_times_interpolant(k):
    v, t = O.unit_box_mesh(1)
    v = v @ np.array([[1.0, 0.2, 0.0], [0.0, 0.9, 0.1], [0.3, 0.0, 1.1]])  # sheared cube
    f = np.array([0.7, -1.3, 0.4])
    b = O.nedelec_vector_load(k, v, t, lambda x, y, z: f)
    m = O.nedelec_csc(k, v, t, alpha=0.0, beta=1.0).to_dense()
    d = np.linalg.solve(m, b)
    rng = np.random.default_rng(k)
    for cell in range(t.shape[0]):
        for _ in range(3):
            xi = rng.dirichlet(np.ones(4))[1:]
            u = O.nedelec_field(k, v, t, d, cell, xi)
            assert np.allclose(u, f, rtol=0, atol=1e-10), (k, cell, u)


@pytest.mark.parametrize("k", [1, 2])
def test_load_is_linear_in_the_source(k):
    v, t = O.unit_box_mesh(1)
    f1 = lambda x, y, z: (np.sin(x), y * y, x * z)
    f2 = lambda x, y, z: (1.0 + z, -x, np.cos(y))
    b1 = O.nedelec_vector_load(k, v, t, f1)
    b2 = O.nedelec_vector_load(k, v, t, f2)
    b3 = O.nedelec_vector_load(k, v, t, lambda x, y, z: tuple(2.0 * a - 3.0 * c for a, c in zip(f1(x, y, z), f2(x, y, z))))
    assert np.allclose(b3, 2.0 * b1 - 3.0 * b2, rtol=0, atol=1e-13)
    assert not np.allclose(b1, 0.0)
    b0 = O.nedelec_vector_load(k, v, t, lambda x, y, z: (0.0, 0.0, 0.0))
    assert np.all(b0 == 0.0)
