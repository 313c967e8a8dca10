"""The colour-conflict validator itself (CPU): it accepts a proper colouring
and rejects one with a planted conflict (two cells sharing a dof in one
colour) or a missing cell."""
import numpy as np
import pytest

from colouring import assert_valid_colouring


class _Fake:
    def __init__(self, gd, cs, cells):
        self.gd, self.cs, self.cells = np.asarray(gd), np.asarray(cs), np.asarray(cells)

    def colouring(self, base=0):
        return self.cs, self.cells + base

    def global_dofs(self, base=0):
        return self.gd + base


def test_validator_accepts_and_rejects():
    # four cells on a strip: cell i holds dofs (i, i + 1); a proper 2-colouring
    gd = [[0, 1], [1, 2], [2, 3], [3, 4]]
    assert assert_valid_colouring(_Fake(gd, [0, 2, 4], [0, 2, 1, 3])) == 2
    with pytest.raises(AssertionError):  # cells 0 and 1 share dof 1 in colour 0
        assert_valid_colouring(_Fake(gd, [0, 2, 4], [0, 1, 2, 3]))
    with pytest.raises(AssertionError):  # cell 3 missing, cell 0 twice
        assert_valid_colouring(_Fake(gd, [0, 2, 4], [0, 2, 1, 0]))
