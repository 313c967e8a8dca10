"""Writes nedelec_high_order.npz: the oracle's Nedelec CSCs at the orders of
the tensor-weighted workload that the CPU oracle is too slow to recompute
inside a GPU test (about 20 s per order-6 element):

* ``two_k{5,6}`` / ``two_k{5,6}w`` -- the reference's two-tetrahedron mesh
  (test/test_tetra_nedelec_sparse_assembly_slow.f90:31-85: vertices 1-5, tets
  [1,2,3,4] and [1,3,2,5], two orientation classes sharing the face 1-2-3),
  curl-mass (alpha = 1.3, beta = 0.7) and tensor-weighted (one SPD tensor per
  cell, beta = 0.7);
* ``box_k4w`` / ``box_k6w`` -- the 1^3 x 6 box with vertex ids relabelled by
  a fixed permutation (six cells in several orientation classes) and seeded
  per-cell SPD tensors K_e = Q diag(exp(0.5 z)) Q^T, beta = 0.6.

The meshes and tensors are stored with the matrices.

    python tests/golden/make_nedelec_high_order.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

TWO_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1.0]])
TWO_T = np.array([[1, 2, 3, 4], [1, 3, 2, 5]], dtype=np.int32)
TWO_K = np.array([[[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]],
                  [[1.1, -0.2, 0.05], [-0.2, 0.8, 0.3], [0.05, 0.3, 1.7]]])


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
    K = np.einsum("tij,tj,tkj->tik", Q, d, Q)
    return 0.5 * (K + K.transpose(0, 2, 1))


def box_mesh():
    import oracle as O

    v, t = O.unit_box_mesh(1)
    perm = np.random.default_rng(20261203).permutation(v.shape[0])
    v2 = np.zeros_like(v)
    v2[perm] = v
    return v2, (perm[t - 1] + 1).astype(np.int32)


def main():
    import oracle as O

    out = {"two_v": TWO_V, "two_t": TWO_T, "two_K": TWO_K}
    bv, bt = box_mesh()
    bK = spd_tensors(bt.shape[0], 3)
    out.update(box_v=bv, box_t=bt, box_K=bK)
    cases = [("two_k5", 5, TWO_V, TWO_T, None, 1.3, 0.7), ("two_k6", 6, TWO_V, TWO_T, None, 1.3, 0.7),
             ("two_k5w", 5, TWO_V, TWO_T, TWO_K, 1.0, 0.7), ("two_k6w", 6, TWO_V, TWO_T, TWO_K, 1.0, 0.7),
             ("box_k4w", 4, bv, bt, bK, 1.0, 0.6), ("box_k6w", 6, bv, bt, bK, 1.0, 0.6)]
    for name, k, v, t, K, alpha, beta in cases:
        a = O.nedelec_csc(k, v, t, alpha=alpha, beta=beta, cell_tensor=K)
        out[name + "_col_ptr"] = a.col_ptr.astype(np.int32)
        out[name + "_row_idx"] = a.row_idx.astype(np.int32)
        out[name + "_val"] = a.val
        print(name, a.nnz, flush=True)
    np.savez_compressed(os.path.join(HERE, "nedelec_high_order.npz"), **out)


if __name__ == "__main__":
    main()
