"""Oracle: Philox4x32-10 and the device initial-mode draws (CPU, NumPy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Philox4x32-10 as published by Salmon, Moraes, Dror and Shaw ("Parallel random
numbers: as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies
(0xD2511F53, 0xCD9E8D57) with Weyl key increments (0x9E3779B9, 0xBB67AE85).
This restates the generator the library's mgk_kmc_init_modes uses (no
reference counterpart: the reference draws initial modes on the host).
"""
import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """ctr: 4 uint32 arrays (broadcastable), key: 2 uint32 scalars -> 4 arrays."""
    c = [np.asarray(x, dtype=np.uint64) for x in ctr]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
        k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
    return [x.astype(np.uint64) for x in c]


def u53(a, b):
    x = (a.astype(np.uint64) << np.uint64(21)) ^ (b.astype(np.uint64) >> np.uint64(11))
    return (x.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def init_modes(nvox, M, seed, mean, std, min_cutoff, max_cutoff, gamma0):
    """(Q0 [M, nvox], cat [2, M, nvox]) as mgk_kmc_init_modes draws them."""
    m = np.repeat(np.arange(M, dtype=np.uint64), nvox)
    v = np.tile(np.arange(nvox, dtype=np.uint64), M)
    key = (seed & MASK, (seed >> 32) & MASK)
    ctr = [m, v & np.uint64(MASK), v >> np.uint64(32), np.zeros_like(m)]
    a = philox4x32_10(ctr, key)
    ctr[3] = np.ones_like(m)
    b = philox4x32_10(ctr, key)

    def bm(x):
        u1, u2 = u53(x[0], x[1]), u53(x[2], x[3])
        r = np.sqrt(-2.0 * np.log(u1))
        return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)
    gxx, gxy = bm(a)
    zq, _ = bm(b)
    q = mean + std * zq
    if min_cutoff is not None or max_cutoff is not None:
        q = np.clip(q, min_cutoff, max_cutoff)
    cat = np.stack([0.5 * gamma0 * gxx, 0.5 * gamma0 * gxy]).reshape(2, M, nvox)
    return q.reshape(M, nvox), cat
