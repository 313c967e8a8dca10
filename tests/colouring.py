"""Colour-conflict validator for the element colourings of the pattern build
(the scatter writes each colour's cells in parallel without atomics, which is
only correct when no two cells of one colour share a dof)."""
import numpy as np


def assert_valid_colouring(p):
    """Every cell appears exactly once, and within each colour no dof is held
    by two cells."""
    cs, cells = p.colouring(base=0)
    gd = p.global_dofs(base=0)
    T = gd.shape[0]
    assert cs[0] == 0 and cs[-1] == T and np.all(np.diff(cs) > 0)
    assert np.array_equal(np.sort(cells), np.arange(T))
    for c in range(cs.shape[0] - 1):
        d = gd[cells[cs[c]:cs[c + 1]]].reshape(-1)
        u = np.unique(d)
        assert u.shape[0] == d.shape[0], f"colour {c}: {d.shape[0] - u.shape[0]} shared dofs"
    return cs.shape[0] - 1
