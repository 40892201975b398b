"""Counter-based right-hand-side generator shared by tests and benchmarks.

value[i] = 2 * u(i) - 1 with u(i) = (splitmix64(seed * 2^32 + i) >> 11) * 2^-53, so any
element can be regenerated from (seed, linear index i) alone, in C, HIP or Python.
Index i is the x-fastest linear index of an (nz, ny, nx) grid.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    z = x + _GAMMA
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform_pm1(shape, seed: int, chunk: int = 1 << 24) -> np.ndarray:
    """U[-1, 1) FP64 array of the given shape from the counter generator."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float64)
    base = np.uint64(seed) << np.uint64(32)
    with np.errstate(over="ignore"):
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            idx = np.arange(s, e, dtype=np.uint64) + base
            z = _splitmix64(idx)
            u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            out[s:e] = 2.0 * u - 1.0
    return out.reshape(shape)
