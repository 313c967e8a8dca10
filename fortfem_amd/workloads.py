"""Synthetic inputs of the benchmark configurations, reproducible bit for bit.

RNG: SplitMix64 stream, u = (x >> 11) * 2^-53, normals by Box-Muller taking
the cosine branch first; seed = 20261203 * 10 + configuration index; per-cell
quantities are drawn in cell order.
"""
from __future__ import annotations

import numpy as np

MASK = (1 << 64) - 1
BASE_SEED = 20261203 * 10


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs (uint64) of a stream started at seed."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = np.maximum(u[0::2], 2.0 ** -53)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


def lognormal_kappa(ncell: int, config: int, sigma: float = 1.0) -> np.ndarray:
    return np.exp(sigma * normal(BASE_SEED + config, ncell))


def uniform_rhs(n: int, config: int, lo: float = -1.0, hi: float = 1.0, offset: int = 7) -> np.ndarray:
    return lo + (hi - lo) * uniform(BASE_SEED + config + offset * 1000003, n)


CONFIGS = {
    "c1_p1_2d": dict(n=65),
    "c2_p2_tet": dict(N=32, degree=2),
    "c4_p1_tet": dict(N=128, degree=1),
}
