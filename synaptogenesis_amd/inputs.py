"""Seeded synthetic inputs shared by the tests, the fixture script and the
benchmark.

Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11): a keyed bijection of a 128-bit counter,
so any element of any stream can be generated independently and in any order.
Workload ``k`` uses key ``(MASTER_SEED, k)`` and counter
``(step, index, stream, 0)``; a uniform in [0,1) is ``(x0 >> 8) * 2**-24``
and a Gaussian is Box-Muller in FP64 over ``x0``/``x1``.
"""

import numpy as np

MASTER_SEED = 20261202

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


_CHUNK = 1 << 18


def _rounds(c0, c1, c2, c3, k0, k1, rounds):
    """Philox rounds over equal-length uint32 arrays, in place."""
    p0 = np.empty(c0.shape, np.uint64)
    p1 = np.empty(c0.shape, np.uint64)
    t = np.empty(c0.shape, np.uint64)
    for r in range(rounds):
        np.multiply(c0, _M0, out=p0, dtype=np.uint64)
        np.multiply(c2, _M1, out=p1, dtype=np.uint64)
        # c0' = hi(p1) ^ c1 ^ k0 ; c2' = hi(p0) ^ c3 ^ k1
        np.right_shift(p1, np.uint64(32), out=t)
        np.bitwise_xor(t.astype(np.uint32), c1, out=c0)
        np.bitwise_xor(c0, k0, out=c0)
        np.right_shift(p0, np.uint64(32), out=t)
        np.bitwise_xor(t.astype(np.uint32), c3, out=c2)
        np.bitwise_xor(c2, k1, out=c2)
        # c1' = lo(p1) ; c3' = lo(p0)
        c1[:] = p1            # uint64 -> uint32 keeps the low word
        c3[:] = p0
        if r != rounds - 1:
            k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32 over uint32 arrays (broadcast together).

    Returns the four output words as uint32 arrays.  Works through the
    arrays in cache-sized chunks.
    """
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32) for c in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    shape = c0.shape
    out = [np.ascontiguousarray(c, dtype=np.uint32).reshape(-1).copy()
           for c in (c0, c1, c2, c3)]
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    n = out[0].size
    with np.errstate(over="ignore"):
        for i in range(0, n, _CHUNK):
            _rounds(*(o[i:i + _CHUNK] for o in out), k0, k1, rounds)
    return tuple(o.reshape(shape) for o in out)


class Stream:
    """One seeded stream: key ``(MASTER_SEED, workload)``."""

    def __init__(self, workload, seed=MASTER_SEED):
        self.k0 = np.uint32(seed & 0xFFFFFFFF)
        self.k1 = np.uint32(workload & 0xFFFFFFFF)

    def words(self, n, step=0, stream=0):
        idx = np.arange(n, dtype=np.uint32)
        return philox4x32(np.uint32(step), idx, np.uint32(stream), 0,
                          self.k0, self.k1)

    def bits(self, n, step=0, stream=0):
        """Raw 32-bit words (uint32)."""
        return self.words(n, step, stream)[0]

    def uniform(self, n, step=0, stream=0, low=0.0, high=1.0):
        """float32 uniforms in [low, high)."""
        x0 = self.words(n, step, stream)[0]
        u = (x0 >> np.uint32(8)).astype(np.float64) * (2.0 ** -24)
        return (low + (high - low) * u).astype(np.float32)

    def integers(self, n, low, high, step=0, stream=0):
        """int32 uniform on [low, high] inclusive."""
        x0 = self.words(n, step, stream)[0].astype(np.uint64)
        span = np.uint64(high - low + 1)
        return (np.int64(low) + (x0 % span).astype(np.int64)).astype(np.int32)

    def normal(self, n, step=0, stream=0, mean=0.0, std=1.0):
        """float32 Gaussians (Box-Muller in FP64)."""
        x0, x1, _, _ = self.words(n, step, stream)
        u1 = ((x0 >> np.uint32(8)).astype(np.float64) + 1.0) * (2.0 ** -24)
        u2 = (x1 >> np.uint32(8)).astype(np.float64) * (2.0 ** -24)
        z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
        return (mean + std * z).astype(np.float32)

    def bernoulli(self, n, p, step=0, stream=0):
        return self.uniform(n, step, stream) < np.float32(p)
