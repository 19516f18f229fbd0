"""Philox4x32-10 counter-based generator and the Box-Muller transform in NumPy.

TEST INFRASTRUCTURE ONLY: the checker of the device SVI sampler
(``gl_svi_diag_sample``).  Restates the published algorithm of Salmon, Moraes,
Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC 2011): ten
rounds of two 32x32 -> 64-bit multiplies, key bumped by the Weyl constants.
Pinned by the authors' known-answer vectors (tests/test_oracle_philox.py).
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32, key: (..., 2) uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint32).copy()
    k1 = np.asarray(key[..., 1], np.uint32).copy()
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ k0.astype(np.uint64), lo1, hi0 ^ c[3] ^ k1.astype(np.uint64), lo0]
        with np.errstate(over="ignore"):
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def unit(u):
    """(0, 1] from 32 random bits: the top 24 bits plus one, times 2^-24."""
    return ((np.asarray(u, np.uint32) >> np.uint32(8)).astype(np.float64) + 1.0) / 16777216.0


def normals4(seed: int, particle, group, step):
    """Four N(0, 1) draws per (particle, group, step) as the device sampler makes them."""
    particle, group, step = np.broadcast_arrays(np.asarray(particle, np.uint32), np.asarray(group, np.uint32),
                                                np.asarray(step, np.uint32))
    ctr = np.stack([particle, group, step, np.zeros_like(particle)], axis=-1)
    key = np.broadcast_to(np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], np.uint32), ctr.shape[:-1] + (2,))
    r = philox4x32_10(ctr, key)
    r0 = np.sqrt(-2.0 * np.log(unit(r[..., 0])))
    r1 = np.sqrt(-2.0 * np.log(unit(r[..., 2])))
    t0 = 2.0 * np.pi * unit(r[..., 1])
    t1 = 2.0 * np.pi * unit(r[..., 3])
    return np.stack([r0 * np.cos(t0), r0 * np.sin(t0), r1 * np.cos(t1), r1 * np.sin(t1)], axis=-1)
