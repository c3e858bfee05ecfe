"""Philox4x32-10 restated in numpy (TEST INFRASTRUCTURE ONLY: the checker of
the start fields of the subspace generation).

  philox4x32_10 ...... J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
                       "Parallel random numbers: as easy as 1, 2, 3", SC11
                       (2011): ten rounds of two 32x32->64-bit products,
                       key bumped by the Weyl constants
  start_field ........ the library's replacement of dfl/dfl_modes.c:248-271
                       (random_fields): component 4 j + i of point x (ipt
                       index) is word i at counter (x, j, 0, 0), j < 6, as
                       (w >> 8) 2^-24 - 1/2; key = two ranlxd numbers of
                       stream 0 times 2^32
Pinned by the known-answer vectors of the Random123 distribution
(tests/test_oracle_philox.py).
"""
import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised over arrays of counters (uint32 arrays or ints)."""
    c = [np.asarray(v, np.uint64) & 0xffffffff for v in (c0, c1, c2, c3)]
    k0 = np.uint64(k0)
    k1 = np.uint64(k1)
    mask = np.uint64(0xffffffff)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & mask, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & mask]
        k0 = (k0 + np.uint64(W0)) & mask
        k1 = (k1 + np.uint64(W1)) & mask
    return [v.astype(np.uint32) for v in c]


def start_field(V, key):
    """(V, 24) float32 start field for the two ranlxd numbers key."""
    k0 = int(key[0] * 4294967296.0)
    k1 = int(key[1] * 4294967296.0)
    x = np.arange(V, dtype=np.uint64)
    out = np.empty((V, 24), np.float32)
    s = np.float32(1.0 / 16777216.0)
    for j in range(6):
        w = philox4x32_10(x, np.full(V, j, np.uint64), 0, 0, k0, k1)
        for i in range(4):
            out[:, 4 * j + i] = (w[i] >> 8).astype(np.float32) * s - np.float32(0.5)
    return out
