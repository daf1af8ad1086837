"""Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC 2011) in numpy, vectorized over counters: the generator
the GPU random configurations use, restated for the tests."""
import numpy as np

M0, M1, W0, W1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57), 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(x, dtype=np.uint64) for x in (c0, c1, c2, c3)]
    k0, k1 = int(k0), int(k1)
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0)) & MASK, p1 & MASK,
             ((p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1)) & MASK, p0 & MASK]
        k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
    return c


def random_qpos(B, lo, hi, seed):
    """(B, nv): entry (b, u) = lo_u + (hi_u - lo_u) * (53-bit fraction of Philox at counter (b, u))."""
    nv = len(lo)
    b = np.repeat(np.arange(B, dtype=np.uint64), nv)
    u = np.tile(np.arange(nv, dtype=np.uint64), B)
    r = philox4x32_10(b & MASK, b >> np.uint64(32), u, np.zeros_like(u), seed & 0xFFFFFFFF, seed >> 32)
    f = ((r[0] >> np.uint64(5)).astype(np.float64) * 67108864.0 + (r[1] >> np.uint64(6)).astype(np.float64)) / 9007199254740992.0
    return (np.asarray(lo)[u.astype(int)] + (np.asarray(hi) - np.asarray(lo))[u.astype(int)] * f).reshape(B, nv)
