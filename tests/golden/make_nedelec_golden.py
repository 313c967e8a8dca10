"""Writes nedelec_single_tet.npz: the oracle's CSC of one tetrahedron at
Nedelec orders 5 and 6 (curl-mass alpha = 1.3, beta = 0.7) and order 5
tensor-weighted, with vertex ids in a non-identity orientation class.  The
GPU parity tests compare against these (the oracle needs ~20 s per order-6
element); tests/test_oracle_nedelec.py re-derives them from the oracle.

    python tests/golden/make_nedelec_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def single_tet():
    """Vertex ids 3, 1, 4, 2 on a perturbed, positively oriented reference tet."""
    ids = np.array([[3, 1, 4, 2]], dtype=np.int32)
    local = np.array([[0.1, -0.05, 0.02], [1.2, 0.1, -0.1], [0.05, 0.9, 0.15], [-0.1, 0.12, 1.1]])
    v = np.zeros((4, 3))
    for l, g in enumerate(ids[0]):
        v[g - 1] = local[l]
    return v, ids


TENSOR = np.array([[[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]]])


def cases():
    v, t = single_tet()
    return [("k5", 5, v, t, None), ("k6", 6, v, t, None), ("k5w", 5, v, t, TENSOR)]


def main():
    import oracle as O

    out = {}
    for name, k, v, t, K in cases():
        a = O.nedelec_csc(k, v, t, alpha=1.3, beta=0.7, cell_tensor=K)
        out[name + "_col_ptr"] = a.col_ptr
        out[name + "_row_idx"] = a.row_idx
        out[name + "_val"] = a.val
    np.savez_compressed(os.path.join(HERE, "nedelec_single_tet.npz"), **out)


if __name__ == "__main__":
    main()
