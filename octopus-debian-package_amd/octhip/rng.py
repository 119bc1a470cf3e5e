"""Counter-based random numbers for the seeded synthetic inputs.

Every random quantity of a workload is a pure function of (seed, stream, counter), so
the same inputs are produced on any host, in any order, without carrying generator
state around.  The mixing function is SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014); normals come from the Box-Muller
transform.

    word(seed, stream, i) = mix64(key(seed, stream) + (i + 1) * GOLDEN)
    key(seed, stream)     = mix64(seed * GOLDEN ^ (stream + 1) * MIX2)
    uniform               = (word >> 11) * 2**-53                in [0, 1)
    normal(i)             = sqrt(-2 ln(1 - u(2i))) * cos(2 pi u(2i+1))
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# default seed of workload k is 0x0C7A0000 + k
SEED_BASE = 0x0C7A0000


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = z.astype(np.uint64, copy=True)
        z ^= z >> np.uint64(30)
        z *= _M1
        z ^= z >> np.uint64(27)
        z *= _M2
        z ^= z >> np.uint64(31)
    return z


def _key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.array([seed], dtype=np.uint64) * GOLDEN
        t = np.array([stream + 1], dtype=np.uint64) * _M2
        return _mix64(s ^ t)[0]


def words(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    k = _key(seed, stream)
    with np.errstate(over="ignore"):
        ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64) * GOLDEN + k
    return _mix64(ctr)


def uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    u = (words(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u = (words(seed, stream, 2 * n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    u1 = 1.0 - u[0::2]          # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def in_ball(seed: int, stream: int, n: int, radius: float) -> np.ndarray:
    """n points uniform in a ball (rejection from the cube, fixed counter budget)."""
    out = np.empty((0, 3))
    chunk = 0
    while out.shape[0] < n:
        c = uniform(seed, stream * 64 + chunk, 3 * 4 * n, -radius, radius).reshape(-1, 3)
        c = c[np.einsum("ij,ij->i", c, c) <= radius * radius]
        out = np.concatenate([out, c])
        chunk += 1
    return out[:n].copy()
