"""GPU parity of mpn_mul_default_mpn_ctx against the oracle and exact ints.

Products are unique, so any exact multiplier pins them: Python's int
product, the CPU oracle (oracle/fsmall_oracle.c, itself pinned to GMP and
Python ints in test_oracle.py), and at full size mod-q checksums.
"""
import sys

import numpy as np
import pytest

from tests.helpers import to_int, rand_limbs

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (2, 1), (3, 3), (7, 5), (16, 16), (64, 1), (100, 37), (255, 256),
          (1000, 999), (1500, 700), (4096, 4096), (5000, 17), (20000, 20000),
          (40000, 40000), (40001, 40001), (65536, 65536), (100000, 3), (131072, 100000)]


@pytest.mark.parametrize("an,bn", SHAPES)
def test_mpn_mul_random(an, bn):
    import flint_amd as fa
    rng = np.random.default_rng(an * 1000003 + bn)
    a = rand_limbs(rng, an)
    b = rand_limbs(rng, bn)
    print("mpn_mul", an, bn, fa.mpn_plan(an, bn), flush=True)
    z = fa.mpn_mul(a, b)
    assert to_int(z) == to_int(a) * to_int(b)


@pytest.mark.parametrize("an,bn", [(1, 1), (33, 33), (3000, 2000), (70000, 70000)])
def test_mpn_mul_all_ones(an, bn):
    """all-ones limbs: long carry-propagate chains through the CRT stitch"""
    import flint_amd as fa
    a = np.full(an, np.uint64(2**64 - 1), dtype=np.uint64)
    b = np.full(bn, np.uint64(2**64 - 1), dtype=np.uint64)
    z = fa.mpn_mul(a, b)
    assert to_int(z) == to_int(a) * to_int(b)


def test_mpn_mul_sparse_and_powers():
    import flint_amd as fa
    for an, bn in [(5000, 5000), (70000, 1000)]:
        a = np.zeros(an, dtype=np.uint64)
        b = np.zeros(bn, dtype=np.uint64)
        a[-1] = 1
        b[-1] = 1
        a[0] = 2**63
        b[bn // 2] = 12345
        z = fa.mpn_mul(a, b)
        assert to_int(z) == to_int(a) * to_int(b)


def test_mpn_mul_matches_oracle():
    import flint_amd as fa
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    for an, bn in [(333, 222), (50000, 45000)]:
        a = rand_limbs(rng, an)
        b = rand_limbs(rng, bn)
        assert np.array_equal(fa.mpn_mul(a, b), O.mpn_mul(a, b))


def test_golden_small_fixtures():
    import os
    import flint_amd as fa
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "mpn_small.npz"))
    for i, (an, bn) in enumerate(f["shapes"]):
        assert np.array_equal(fa.mpn_mul(f["a%d" % i], f["b%d" % i]), f["z%d" % i])


@pytest.mark.parametrize("name", ["config1_65536", "config2_2p24"])
def test_config_checksums(name):
    """full config-size products against the GMP checksums (SHA-256 and
    residues mod three primes), inputs from the seeded generator"""
    import hashlib
    import json
    import os
    import flint_amd as fa
    with open(os.path.join(os.path.dirname(__file__), "golden", "checksums.json")) as fh:
        rec = json.load(fh)[name]
    a, b = fa.synth_mpn_pair(rec["seed"], rec["an"], rec["bn"])
    assert hashlib.sha256(a.tobytes()).hexdigest() == rec["a_sha256"]
    print(name, fa.mpn_plan(rec["an"], rec["bn"]), flush=True)
    z = fa.mpn_mul(a, b)
    assert hashlib.sha256(z.tobytes()).hexdigest() == rec["sha256"]
