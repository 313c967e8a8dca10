"""Writes nedelec_points.npz: the oracle's tensor-weighted Nedelec CSCs with
the coefficient evaluated at every point of the rule
(assemble_tetra_nedelec_weighted_element_order, assembly_tetra_nedelec_3d.f90:
2452-2514), at orders 5 and 6, where the CPU oracle is too slow to recompute
inside a GPU test:

* ``two_k5p`` / ``two_k6p`` -- the reference's two-tetrahedron mesh
  (test/test_tetra_nedelec_sparse_assembly_slow.f90:31-85: two orientation
  classes sharing the face 1-2-3), beta = 0.7, with the symmetric tensor
  field K(x) = K0 + x K1 + y K2 + z^2 K3 (K0 positive definite) sampled at
  the rule's points mapped into each cell.

The samples (T, nq, 3, 3) are stored with the matrices, so the GPU test
feeds exactly these inputs.

    python tests/golden/make_nedelec_points.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

TWO_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1.0]])
TWO_T = np.array([[1, 2, 3, 4], [1, 3, 2, 5]], dtype=np.int32)
K0 = np.array([[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]])
K1 = np.array([[0.4, -0.1, 0.2], [-0.1, 0.3, 0.0], [0.2, 0.0, -0.2]])
K2 = np.array([[-0.3, 0.2, 0.1], [0.2, 0.5, -0.1], [0.1, -0.1, 0.2]])
K3 = np.array([[0.2, 0.0, 0.1], [0.0, -0.1, 0.05], [0.1, 0.05, 0.3]])
BETA = 0.7


def samples(k, v, t):
    import oracle as O

    pts = O.nedelec_rule_points(k)
    X = O.mapped_points(v, t, pts)
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    return (K0 + x[..., None, None] * K1 + y[..., None, None] * K2
            + (z * z)[..., None, None] * K3)


def main():
    import oracle as O

    out = {"two_v": TWO_V, "two_t": TWO_T, "beta": np.array(BETA)}
    for k in (5, 6):
        Kp = samples(k, TWO_V, TWO_T)
        a = O.nedelec_csc_points(k, TWO_V, TWO_T, Kp, beta=BETA)
        name = f"two_k{k}p"
        out[name + "_K"] = Kp
        out[name + "_col_ptr"] = a.col_ptr
        out[name + "_row_idx"] = a.row_idx
        out[name + "_val"] = a.val
        print(name, a.n_rows, a.nnz, flush=True)
    np.savez_compressed(os.path.join(HERE, "nedelec_points.npz"), **out)


if __name__ == "__main__":
    main()
