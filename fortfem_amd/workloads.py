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


def served_stream(count: int, config: int = 5):
    """The served request stream (configuration 5): request i draws, from its
    own seed BASE_SEED + config + 1000 i, a type (P1 35 %, P2 35 %, Nedelec
    k=1 15 %, k=2 15 %), the cube size N in 8..48 and 1..8 right-hand sides.
    Returns a list of dicts (family code 1 = H1, 2 = Nedelec)."""
    out = []
    for i in range(count):
        u = uniform(BASE_SEED + config + 1000 * i, 3)
        t = u[0]
        if t < 0.35:
            fam, k = 1, 1
        elif t < 0.70:
            fam, k = 1, 2
        elif t < 0.85:
            fam, k = 2, 1
        else:
            fam, k = 2, 2
        N = 8 + int(u[1] * 41)
        nrhs = 1 + int(u[2] * 8)
        out.append({"index": i, "family": fam, "order": k, "N": N, "nrhs": nrhs,
                    "seed": BASE_SEED + config + 1000 * i})
    return out


def request_data(req: dict, ncell: int, n_free: int):
    """Per-cell lognormal coefficient and U[-1, 1] right-hand sides (nrhs, n_free)."""
    coeff = np.exp(normal(req["seed"] + 1, ncell))
    rhs = -1.0 + 2.0 * uniform(req["seed"] + 2, req["nrhs"] * n_free).reshape(req["nrhs"], n_free)
    return coeff, rhs

