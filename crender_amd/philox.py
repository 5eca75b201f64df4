"""Vectorised Philox4x32-10 (Salmon et al., SC'11) in numpy.

Used by the scene generators so that every synthetic scene is a pure function of its
seed.  The same counter layout conventions as the device RNG apply: key = (seed,
0x43524E44); generator streams use counter word 2 = 0xFFFFFF00 | stream, which never
collides with the render dimensions (bounce index or 0xFFFFFFFF).
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
KEY_HI = 0x43524E44
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Returns the four output words for counters (c0..c3) and key (k0, k1)."""
    c0 = np.asarray(c0, dtype=np.uint32)
    c1 = np.asarray(c1, dtype=np.uint32)
    c2 = np.asarray(c2, dtype=np.uint32)
    c3 = np.asarray(c3, dtype=np.uint32)
    shape = np.broadcast_shapes(c0.shape, c1.shape, c2.shape, c3.shape)
    c0, c1, c2, c3 = (np.broadcast_to(x, shape).astype(np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
                k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def u01(x) -> np.ndarray:
    """u32 -> float32 in [0, 1): (x >> 8) * 2^-24."""
    return ((np.asarray(x, dtype=np.uint32) >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0))


class Stream:
    """Counter-based stream of uniforms for one generator purpose."""

    def __init__(self, seed: int, stream: int):
        self.k0 = int(seed) & 0xFFFFFFFF
        self.k1 = KEY_HI
        self.word2 = 0xFFFFFF00 | (int(stream) & 0xFF)

    def uniform(self, n: int, block: int = 0, dims: int = 4) -> np.ndarray:
        """(n, dims) float32 uniforms; element (i, d) uses counter (i, d//4, word2, block)."""
        i = np.arange(n, dtype=np.uint32)
        cols = []
        for j in range((dims + 3) // 4):
            out = philox4x32_10(i, np.uint32(j), np.uint32(self.word2), np.uint32(block), self.k0, self.k1)
            cols.extend(out)
        return np.stack([u01(c) for c in cols[:dims]], axis=1)

    def normal(self, n: int, block: int = 0, dims: int = 3) -> np.ndarray:
        """(n, dims) standard normals by Box-Muller (float64 internally)."""
        u = self.uniform(n, block=block, dims=2 * ((dims + 1) // 2)).astype(np.float64)
        u1 = np.maximum(u[:, 0::2], 1.0 / 16777216.0)
        u2 = u[:, 1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)], axis=1)
        return z[:, :dims]
