"""Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, SC'11).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
Pinned by the published known-answer vectors of the Random123 distribution
(kat_vectors, philox4x32 with 10 rounds), checked in
tests/test_oracle_golden.py::test_philox_known_answers.
"""
from __future__ import annotations

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """ctr (n,4) uint32, key (n,2) uint32 -> (n,4) uint32."""
    c = np.asarray(ctr, dtype=np.uint64) & MASK
    k = np.asarray(key, dtype=np.uint64) & MASK
    c0, c1, c2, c3 = (c[:, i].copy() for i in range(4))
    k0, k1 = k[:, 0].copy(), k[:, 1].copy()
    for r in range(10):
        p0 = c0 * np.uint64(M0)
        p1 = c2 * np.uint64(M1)
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & np.uint64(MASK), lo1, (hi0 ^ c3 ^ k1) & np.uint64(MASK), lo0
        if r < 9:
            k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
            k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
    return np.stack([c0, c1, c2, c3], axis=1).astype(np.uint32)


def to_uniform(x):
    """32-bit word -> float32-exact uniform in (0, 1]: ((x >> 8) + 1) * 2^-24."""
    return ((np.asarray(x, dtype=np.uint64) >> np.uint64(8)) + np.uint64(1)).astype(np.float64) * 2.0 ** -24
