"""Shared helpers for the parity tests: seeded inputs and exact references."""
import numpy as np


def to_int(a):
    """little-endian uint64 limbs -> Python int"""
    a = np.asarray(a, dtype=np.uint64)
    return int.from_bytes(a.tobytes(), "little")


def from_int(x, n):
    return np.frombuffer(int(x).to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def rand_limbs(rng, n, top_nonzero=True):
    a = rng.integers(0, 2**64, size=n, dtype=np.uint64, endpoint=False)
    if top_nonzero and n:
        while a[-1] == 0:
            a[-1] = rng.integers(1, 2**64, dtype=np.uint64, endpoint=False)
    return a
