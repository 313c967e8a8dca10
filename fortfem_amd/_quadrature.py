"""Host-side quadrature/basis helpers used to marshal callback arguments.

The collapsed rule restates tetra_duffy_quadrature
(src/quadrature/tetra_duffy_quadrature.f90:12-51) with ascending
Gauss-Legendre nodes, the order the HIP library tabulates; it only serves to
place the points at which a Python source callback is sampled.
"""
from __future__ import annotations

import numpy as np


def gauss_legendre01(m: int):
    x, w = np.polynomial.legendre.leggauss(m)
    return 0.5 * (x + 1.0), 0.5 * w


def tetra_duffy(p: int):
    if p < 0:
        raise ValueError("polynomial degree must be >= 0")
    m = (p + 4) // 2
    s, sw = gauss_legendre01(m)
    a, b, c = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    a, b, c = a.ravel(), b.ravel(), c.ravel()
    c1 = 1.0 - s[a]
    c2 = 1.0 - s[b]
    x = s[a]
    y = c1 * s[b]
    z = c1 * c2 * s[c]
    w = sw[a] * sw[b] * sw[c] * (c1 * c1) * c2
    return x, y, z, w


def lagrange_nodes(d: int) -> np.ndarray:
    """initialize_tetra_lagrange node order (tetra_lagrange_arbitrary_order.f90:51-61)."""
    out = []
    for a in range(d + 1):
        for b in range(d + 1 - a):
            for c in range(d + 1 - a - b):
                out.append((b / d, c / d, (d - a - b - c) / d))
    return np.array(out)
