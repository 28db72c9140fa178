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


@pytest.mark.parametrize("n", [1 << 16, 1 << 20, 3 << 20])
def test_mpn_mul_all_ones_closed_form(n):
    """(2^(64n) - 1)^2 = 2^(128n) - 2^(64n+1) + 1: limb 0 is 1, limbs 1..n-1
    are 0, limb n is 2^64 - 2, the rest all ones.  Every CRT workgroup
    propagates, so the carry resolution crosses the whole product."""
    import flint_amd as fa
    a = np.full(n, np.uint64(2**64 - 1), dtype=np.uint64)
    z = fa.mpn_mul(a, a)
    want = np.full(2 * n, np.uint64(2**64 - 1), dtype=np.uint64)
    want[0] = 1
    want[1:n] = 0
    want[n] = np.uint64(2**64 - 2)
    assert np.array_equal(z, want)


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


@pytest.mark.parametrize("an,bn,np_,bits", [(65536, 65536, 6, 136), (60000, 50000, 6, 136),
                                           (1 << 20, 1 << 20, 6, 136), (1000000, 3, 4, 88),
                                           (700000, 650000, 8, 184), (3 << 20, 2 << 20, 6, 136),
                                           (30000, 30000, 3, 64), (5000, 4000, 2, 40),
                                           (20000000, 100, 6, 136)])
def test_fast_kernels_match_generic(an, bn, np_, bits):
    """the register radix-16 kernels and the generic LDS kernels give the
    same product (and both the exact one)"""
    import flint_amd as fa
    rng = np.random.default_rng(an ^ bn)
    a = rand_limbs(rng, an)
    b = rand_limbs(rng, bn)
    print("fast-vs-generic", an, bn, np_, bits, flush=True)
    try:
        fa.set_fast_kernels(False)
        z0 = fa.mpn_mul_plan(a, b, np_, bits)
    finally:
        fa.set_fast_kernels(True)
    z1 = fa.mpn_mul_plan(a, b, np_, bits)
    assert np.array_equal(z0, z1)
    if an * bn < (1 << 34):
        assert to_int(z1) == to_int(a) * to_int(b)
    else:
        q = (1 << 61) - 1
        assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z1) % q


@pytest.mark.parametrize("depth", list(range(13, 25)))
def test_fast_kernels_every_depth(depth):
    """every column depth d1 = depth - 12 of the radix-16 kernels against the
    generic kernels and an exact check, with the plan forced to np=6/136"""
    import flint_amd as fa
    n = max(1, int(0.7 * 2 ** depth * 136 / 128))
    rng = np.random.default_rng(depth)
    a = rand_limbs(rng, n)
    b = rand_limbs(rng, n - 7 if n > 8 else n)
    z1 = fa.mpn_mul_plan(a, b, 6, 136)
    try:
        fa.set_fast_kernels(False)
        z0 = fa.mpn_mul_plan(a, b, 6, 136)
    finally:
        fa.set_fast_kernels(True)
    assert np.array_equal(z0, z1)
    q = (1 << 61) - 1
    assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z1) % q


@pytest.mark.parametrize("an,bn", [(1 << 24, 1 << 24), (15000000, 14000000)])
def test_column_tile_variants_agree(an, bn):
    """8-column, 4-column half-LDS and 4-column whole-LDS tiles of the
    2^24-point column passes give the same product"""
    import flint_amd as fa
    a, b = fa.synth_mpn_pair(0xC01, an, bn)
    print("col-variants", an, bn, fa.mpn_plan(an, bn), flush=True)
    z = {}
    try:
        for v in (0, 2, 1):
            fa.set_col_variant(v)
            z[v] = fa.mpn_mul(a, b)
    finally:
        fa.set_col_variant(1)
    z1 = z[1]
    assert np.array_equal(z[0], z1)
    assert np.array_equal(z[2], z1)
    q = (1 << 61) - 1
    assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z1) % q


@pytest.mark.parametrize("an,bn", [(1 << 24, 1 << 24), (14000000, 16000000), (30000000, 2000000),
                                   (1 << 24, 9000000)])
def test_prime_groups_agree(an, bn):
    """the 2^24-point passes in one launch per pass, and with the primes in two
    or three groups on forked streams, give the same product (plans with 6
    and 5 primes, an operand longer than half the transform); the device
    entry point on a caller's stream included"""
    import torch
    import flint_amd as fa
    a, b = fa.synth_mpn_pair(0xC03, an, bn)
    plan = fa.mpn_plan(an, bn)
    print("prime-groups", an, bn, plan, flush=True)
    assert plan["depth"] == 24
    z = {}
    try:
        for v in (0, 3, 2):
            fa.set_prime_groups(v)
            z[v] = fa.mpn_mul(a, b)
        dev = torch.device("cuda:0")
        ta = torch.from_numpy(a.view(np.int64)).to(dev)
        tb = torch.from_numpy(b.view(np.int64)).to(dev)
        tz = torch.zeros(an + bn, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()   # the fill runs on torch's stream, the product on another
        stream = torch.cuda.Stream()
        fa.mpn_mul_device(tz.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn, stream.cuda_stream)
        stream.synchronize()
        zd = tz.cpu().numpy().view(np.uint64)
    finally:
        fa.set_prime_groups(2)
    assert np.array_equal(z[0], z[2])
    assert np.array_equal(z[0], z[3])
    assert np.array_equal(z[0], zd)
    q = (1 << 61) - 1
    assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z[0]) % q


@pytest.mark.parametrize("an,bn", [(1 << 24, 1 << 24), (3000000, 2500000)])
def test_row_variants_agree(an, bn):
    """the row pass with all last-stage twiddles held and with them formed
    per radix-4 layer give the same product"""
    import flint_amd as fa
    a, b = fa.synth_mpn_pair(0xC02, an, bn)
    print("row-variants", an, bn, fa.mpn_plan(an, bn), flush=True)
    z = {}
    try:
        for v in (1, 0):
            fa.set_row_variant(v)
            z[v] = fa.mpn_mul(a, b)
    finally:
        fa.set_row_variant(1)
    assert np.array_equal(z[0], z[1])
    q = (1 << 61) - 1
    assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z[0]) % q


@pytest.mark.parametrize("an,bn,np_,bits", [(65536, 65536, 6, 136), (5000, 4000, 2, 40),
                                           (30000, 30000, 3, 64), (1000000, 3, 4, 88),
                                           (200000, 190000, 5, 112), (700000, 650000, 8, 184),
                                           (300000, 300000, 7, 160), (1 << 20, 1 << 20, 6, 136)])
def test_crt_variants_agree(an, bn, np_, bits):
    """FP64-digit CRT and 64-bit-limb CRT give the same (exact) product for
    every prime count"""
    import flint_amd as fa
    rng = np.random.default_rng(an + 7 * bn + np_)
    a = rand_limbs(rng, an)
    b = rand_limbs(rng, bn)
    print("crt-variants", an, bn, np_, bits, flush=True)
    try:
        fa.set_crt_variant(0)
        z0 = fa.mpn_mul_plan(a, b, np_, bits)
    finally:
        fa.set_crt_variant(1)
    z1 = fa.mpn_mul_plan(a, b, np_, bits)
    assert np.array_equal(z0, z1)
    q = (1 << 61) - 1
    assert (to_int(a) % q) * (to_int(b) % q) % q == to_int(z1) % q


def test_every_plan_of_the_2p24_product_agrees():
    """the four valid plans of 2^24 x 2^24 limbs (6 primes at 136 and 128
    bits, 7 at 160, 8 at 184; every one a full 2^24-point transform) give
    the same product, which matches the residue check mod 2^61 - 1"""
    import flint_amd as fa
    n = 1 << 24
    a, b = fa.synth_mpn_pair(0x2424, n, n)
    q = (1 << 61) - 1
    want = (to_int(a) % q) * (to_int(b) % q) % q
    z0 = None
    for np_, bits in ((6, 136), (6, 128), (7, 160), (8, 184)):
        print("plan-agree", np_, bits, flush=True)
        z = fa.mpn_mul_plan(a, b, np_, bits)
        if z0 is None:
            z0 = z
            assert to_int(z) % q == want
        else:
            assert np.array_equal(z, z0), (np_, bits)


@pytest.mark.parametrize("an,bn,cap", [(2000, 1500, 3000), (5000, 300, 3000), (4096, 4096, 3000),
                                       (7000, 2000, 2500)])
def test_blocked_product_exact(an, bn, cap):
    """integer products in block products added on the device (forced by a
    small block cap) against exact integers; all-ones operands make the
    carries ripple across blocks"""
    import flint_amd as fa
    rng = np.random.default_rng(an * 3 + bn)
    cases = [(rand_limbs(rng, an), rand_limbs(rng, bn)),
             (np.full(an, ~np.uint64(0), dtype=np.uint64), np.full(bn, ~np.uint64(0), dtype=np.uint64))]
    print("blocked", an, bn, cap, flush=True)
    try:
        fa.set_mpn_block_cap(cap)
        for a, b in cases:
            z = fa.mpn_mul(a, b)
            want = to_int(a) * to_int(b)
            assert to_int(z) == want
    finally:
        fa.set_mpn_block_cap(0)


def test_blocked_matches_single_plan_on_device_stream():
    import torch
    import flint_amd as fa
    an, bn = 200000, 150000
    a, b = fa.synth_mpn_pair(0xB10C, an, bn)
    z0 = fa.mpn_mul(a, b)
    dev = torch.device("cuda:0")
    ta = torch.from_numpy(a.view(np.int64)).to(dev)
    tb = torch.from_numpy(b.view(np.int64)).to(dev)
    tz = torch.zeros(an + bn, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()   # the fill runs on torch's stream, the product on another
    stream = torch.cuda.Stream()
    try:
        fa.set_mpn_block_cap(120000)
        fa.mpn_mul_device(tz.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn, stream.cuda_stream)
        stream.synchronize()
    finally:
        fa.set_mpn_block_cap(0)
    assert np.array_equal(tz.cpu().numpy().view(np.uint64), z0)


def test_product_beyond_one_plan():
    """2^25 x 2^25 limbs: no single 2^24-point plan covers it; the block
    path computes it (checked mod three 61-bit primes)"""
    import flint_amd as fa
    n = 1 << 25
    with pytest.raises(fa.FlintAmdError):
        fa.mpn_plan(n, n)
    a, b = fa.synth_mpn_pair(0xB1C, n, n)
    print("beyond-plan", n, flush=True)
    z = fa.mpn_mul(a, b)
    A, B, Z = to_int(a), to_int(b), to_int(z)
    for q in ((1 << 61) - 1, (1 << 61) - 31, (1 << 61) - 45):
        assert (A % q) * (B % q) % q == Z % q
