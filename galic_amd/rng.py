"""Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC 2011).

The reference draws from GSL's ranlxd1 stream seeded 42+rank (init.c:25-27),
so its exact sequence cannot be reproduced without GSL.  The build instead
keys every draw by (seed, stream) and counts by (item id, attempt, draw), so
any draw can be regenerated independently on the host or on the GPU.
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32: counters c0..c3 (uint32 arrays), key (k0, k1)."""
    c0 = np.asarray(c0, dtype=np.uint32).copy()
    c1 = np.asarray(c1, dtype=np.uint32).copy()
    c2 = np.asarray(c2, dtype=np.uint32).copy()
    c3 = np.asarray(c3, dtype=np.uint32).copy()
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(rounds):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & _MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + _W0)
            k1 = np.uint32(k1 + _W1)
    return c0, c1, c2, c3


def _u01(a, b):
    """Double in (0,1) from two 32-bit words (53 random bits, never 0 or 1)."""
    hi = (a >> np.uint32(5)).astype(np.float64)   # 27 bits
    lo = (b >> np.uint32(6)).astype(np.float64)   # 26 bits
    return (hi * 67108864.0 + lo + 0.5) / 9007199254740992.0


def uniforms(seed, stream, ids, attempt=0, block=0):
    """Four independent uniforms per id: shape (4, len(ids))."""
    ids = np.asarray(ids, dtype=np.uint64)
    lo = (ids & _MASK).astype(np.uint32)
    hi = (ids >> np.uint64(32)).astype(np.uint32)
    out = []
    for half in range(2):
        r = philox4x32(lo, hi, np.full_like(lo, attempt), np.full_like(lo, 2 * block + half),
                       np.uint32(seed & 0xFFFFFFFF), np.uint32(stream & 0xFFFFFFFF))
        out.append(_u01(r[0], r[1]))
        out.append(_u01(r[2], r[3]))
    return np.stack(out)


def gaussians(seed, stream, ids, attempt=0, block=0):
    """Four independent standard normals per id (Box-Muller): shape (4, len(ids))."""
    u = uniforms(seed, stream, ids, attempt, block)
    r1 = np.sqrt(-2.0 * np.log(u[0]))
    r2 = np.sqrt(-2.0 * np.log(u[2]))
    return np.stack([r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]),
                     r2 * np.cos(2 * np.pi * u[3]), r2 * np.sin(2 * np.pi * u[3])])
