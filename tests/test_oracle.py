"""CPU tests pinning the oracle (oracle/fsmall_oracle.c).

Each check compares the oracle with something it does not share code with:
tables printed in the reference's own comments, GMP's mpn_mul (the
multiplier FLINT's t-mul.c:54 tests against), Python integers, brute-force
evaluation and the closed-form transform invariants of t-sd_fft.c.
"""
import ctypes
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import to_int
from tests.golden.make_fixtures import gmp, gmp_mul, poly_ks, rng_pair

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# mpn_mul.c:71-78
REF_PRIMES = [1108307720798209, 659706976665601, 1086317488242689, 910395627798529,
              699289395265537, 1022545813831681, 1013749720809473, 868614185943041]
# mpn_mul.c:105-119
REF_BN_BOUNDS = {(4, 84): 2536637511, (4, 88): 10380583, (4, 92): 42390,
                 (5, 112): 32822700, (5, 116): 132790, (5, 120): 534,
                 (6, 136): 144789875, (6, 140): 582218, (6, 144): 2337,
                 (7, 160): 613493872, (7, 164): 2456369, (7, 168): 9826,
                 (8, 184): 2177184315, (8, 188): 8689506, (8, 192): 34662}


def test_default_primes_match_reference_table():
    assert O.default_primes() == REF_PRIMES
    for p in REF_PRIMES:
        assert O.lib().or_is_prime(p)


def test_bn_bounds_match_reference_table():
    for (np_, bits), bound in REF_BN_BOUNDS.items():
        assert O.lib().or_crt_bn_bound(np_, bits) == bound, (np_, bits)


def test_prime_products_log2():
    # log2 of p_0 ... p_(k-1) for the default primes, to two decimals
    import math
    expect = [49.98, 99.21, 149.15, 198.85, 248.16, 298.02, 347.87, 397.50]
    acc = 1
    for k, p in enumerate(REF_PRIMES):
        acc *= p
        assert abs(math.log2(acc) - expect[k]) < 0.01


def test_is_prime_against_trial_division():
    rnd = random.Random(3)
    for _ in range(300):
        n = rnd.randrange(2, 10**6)
        tp = n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1))
        assert bool(O.lib().or_is_prime(n)) == tp


def test_mulmod_satisfies_bounds():
    L = O.lib()
    for p in REF_PRIMES + [257, 0x0003f00000000001, 7340033]:
        assert L.or_mulmod_satisfies_bounds(p)
    # 53-bit primes leave no room for the exact product split
    assert not L.or_mulmod_satisfies_bounds((1 << 53) - 111)


@pytest.mark.parametrize("p", [REF_PRIMES[0], REF_PRIMES[1], 7340033, 257])
def test_fp_mulmod_exact(p):
    """machine_vectors.h:681-704: products in (-2n^2, 2n^2) land in (-n, n),
    products in (-4n^2, 4n^2) in (-3n/2, 3n/2), congruent to a*b mod n"""
    L = O.lib()
    rnd = random.Random(p)
    n = float(p)
    ninv = 1.0 / n
    for _ in range(4000):
        a = rnd.randrange(-2 * p + 1, 2 * p)
        bound2 = 2 * p * p
        b = rnd.randrange(-p // 2, p // 2 + 1)
        r = L.or_fp_mulmod(float(a), float(b), n, ninv)
        assert r == int(r)
        r = int(r)
        assert (r - a * b) % p == 0
        if abs(a * b) < bound2:
            assert abs(r) < p
        a4 = rnd.randrange(-3 * p, 3 * p)
        b4 = rnd.randrange(-(4 * p * p) // (3 * p), (4 * p * p) // (3 * p))
        if abs(a4 * b4) < 4 * p * p:
            r4 = int(L.or_fp_mulmod(float(a4), float(b4), n, ninv))
            assert (r4 - a4 * b4) % p == 0
            assert 2 * abs(r4) < 3 * p
    for _ in range(2000):
        a = rnd.randrange(-(1 << 52), 1 << 52)
        r = int(L.or_fp_reduce_to_0n(float(a), n, ninv))
        assert r == a % p


def test_trunc_index_involution():
    for L_ in (0, 3, 4, 10):
        perm = [O.lib().or_trunc_index(L_, i) for i in range(1 << L_)]
        assert sorted(perm) == list(range(1 << L_))
        assert all(perm[perm[i]] == i for i in range(1 << L_))
    # 4x4 transposition inside 16-blocks (fft_small.h:247-253)
    assert O.lib().or_trunc_index(4, 1) == 4
    assert O.lib().or_trunc_index(8, 16 + 6) == 16 + 9


@pytest.mark.parametrize("p", [0x0003f00000000001, 257, REF_PRIMES[3]])
def test_root_table(p):
    """w[j] = e(theta(j)): w[0]=1, w[1]=-1, w[2j]^2=w[j], w[2j+1]=-w[2j]"""
    L = O.lib()
    v = ((p - 1) & -(p - 1)).bit_length() - 1
    assert L.or_w(p, 0) == 1
    assert L.or_w(p, 1) == p - 1
    g = L.or_primitive_2power_root(p)
    assert pow(g, 1 << v, p) == 1 and pow(g, 1 << (v - 1), p) == p - 1
    for j in list(range(1, 64)) + [12345 % (1 << (v - 1))]:
        if 2 * j + 1 >= 1 << v:
            continue
        w2j = L.or_w(p, 2 * j)
        assert pow(w2j, 2, p) == L.or_w(p, j)
        assert L.or_w(p, 2 * j + 1) == (p - w2j) % p


def horner(X, n, pt, p):
    r = 0
    for c in reversed(X[:n]):
        r = (r * pt + int(c)) % p
    return r


@pytest.mark.parametrize("p,maxL", [(0x0003f00000000001, 13), (257, 8)])
def test_fft_trunc_matches_evaluation(p, maxL):
    """t-sd_fft.c:45-66 on the oracle"""
    rnd = random.Random(11)
    for L_ in range(0, maxL + 1):
        N = 1 << L_
        for rep in range(4):
            X = [rnd.randrange(p) for _ in range(N)]
            it = N if rep == 0 else rnd.randint(1, N)
            ot = N if rep == 0 else rnd.randint(1, N)
            d = np.zeros(N, dtype=np.uint64)
            d[:it] = X[:it]
            out = O.fft_trunc(p, d, L_, it, ot)
            for _ in range(6):
                i = rnd.randrange(ot)
                assert int(out[O.lib().or_trunc_index(L_, i)]) == \
                    horner(X, it, O.lib().or_w(p, i), p)


@pytest.mark.parametrize("p,maxL", [(0x0003f00000000001, 13), (257, 8)])
def test_ifft_trunc_roundtrip(p, maxL):
    """t-sd_fft.c:68-88 on the oracle: ifft(fft(x)) = 2^L x"""
    rnd = random.Random(12)
    for L_ in range(0, maxL + 1):
        N = 1 << L_
        for rep in range(4):
            tr = N if rep == 0 else rnd.randint(1, N)
            X = [rnd.randrange(p) for _ in range(tr)]
            d = np.zeros(N, dtype=np.uint64)
            d[:tr] = X
            y = O.ifft_trunc(p, O.fft_trunc(p, d, L_, tr, tr), L_, tr)
            m = pow(2, L_, p)
            assert [int(v) for v in y[:tr]] == [(x * m) % p for x in X]


def test_ifft_from_brute_force_values():
    """the inverse recovers coefficients from values computed by brute force"""
    p = REF_PRIMES[0]
    rnd = random.Random(5)
    for L_ in (4, 9, 10):
        N = 1 << L_
        tr = rnd.randint(N // 2 + 1, N) if L_ > 8 else N
        X = [rnd.randrange(p) for _ in range(tr)]
        vals = np.zeros(N, dtype=np.uint64)
        for i in range(((tr + 255) // 256) * 256 if L_ > 8 else N):
            vals[O.lib().or_trunc_index(L_, i)] = horner(X, tr, O.lib().or_w(p, i), p)
        y = O.ifft_trunc(p, vals, L_, tr)
        m = pow(2, L_, p)
        assert [int(v) for v in y[:tr]] == [(x * m) % p for x in X]


def test_mpn_to_residues_closed_form():
    """t-mpn_ctx_to_from_ffts.c: chunk i = bits [i*bits, (i+1)*bits) mod p"""
    rnd = np.random.default_rng(4)
    a = rnd.integers(0, 2**64, 50, dtype=np.uint64)
    A = to_int(a)
    for bits in (40, 45, 64, 84, 88, 136, 140, 184, 192):
        for p in REF_PRIMES[:3]:
            dn = (64 * 50 + bits - 1) // bits + 3
            d = O.mpn_to_residues(p, a, bits, dn)
            for i in range(dn):
                assert int(d[i]) == ((A >> (i * bits)) & ((1 << bits) - 1)) % p


def test_plans_at_the_benchmark_shapes():
    """mpn_ctx_best_profile / _fft_small_np2_choose at the benchmark shapes
    (2^16 and 2^24 limbs on 1 and 8 threads) and at the two-prime limit"""
    assert O.mpn_plan(65536, 65536, 1) == dict(np=6, bits=140, depth=16, ztrunc=60160,
                                               alen=29960, blen=29960)
    assert O.mpn_plan(65536, 65536, 8)["np"] == 8
    assert O.mpn_plan(65536, 65536, 8)["ztrunc"] == 44800
    assert O.mpn_plan(1 << 24, 1 << 24, 1) == dict(np=6, bits=136, depth=24, ztrunc=15790336,
                                                   alen=7895161, blen=7895161)
    p8 = O.mpn_plan(1 << 24, 1 << 24, 8)
    assert (p8["np"], p8["bits"], p8["ztrunc"]) == (8, 184, 11671296)
    assert O.mpn_plan(2000, 1500)["np"] == 2
    assert O.mpn_plan(40000, 40000)["np"] == 2
    assert O.mpn_plan(40001, 40001)["np"] >= 4


@pytest.mark.skipif(gmp() is None, reason="libgmp not present")
@pytest.mark.parametrize("an,bn", [(1, 1), (3, 2), (100, 37), (2000, 1500), (39999, 40000),
                                   (40001, 2), (50000, 3), (70000, 65536)])
def test_oracle_mpn_mul_vs_gmp(an, bn):
    rng = np.random.default_rng(an + 7 * bn)
    a = rng.integers(0, 2**64, an, dtype=np.uint64)
    b = rng.integers(0, 2**64, bn, dtype=np.uint64)
    a[-1] |= np.uint64(1)
    b[-1] |= np.uint64(1)
    z = O.mpn_mul(a, b)
    assert np.array_equal(z, gmp_mul(a, b))
    assert to_int(z) == to_int(a) * to_int(b)


def test_oracle_mpn_mul_all_ones():
    for an, bn in [(41000, 41000), (1000, 999)]:
        a = np.full(an, np.uint64(2**64 - 1))
        b = np.full(bn, np.uint64(2**64 - 1))
        assert to_int(O.mpn_mul(a, b)) == to_int(a) * to_int(b)


def test_golden_small_mpn_vs_oracle():
    f = np.load(os.path.join(GOLDEN, "mpn_small.npz"))
    for i, (an, bn) in enumerate(f["shapes"]):
        a, b = rng_pair(0xF1A70000 + 100 + i, int(an), int(bn))
        assert np.array_equal(a, f["a%d" % i]) and np.array_equal(b, f["b%d" % i])
        assert np.array_equal(O.mpn_mul(a, b), f["z%d" % i])


def test_golden_config1_checksum_vs_oracle():
    import hashlib
    with open(os.path.join(GOLDEN, "checksums.json")) as fh:
        rec = json.load(fh)["config1_65536"]
    a, b = rng_pair(rec["seed"], rec["an"], rec["bn"])
    z = O.mpn_mul(a, b)
    assert hashlib.sha256(z.tobytes()).hexdigest() == rec["sha256"]


def test_nmod_oracle_vs_fixtures():
    f = np.load(os.path.join(GOLDEN, "poly_small.npz"))
    for k, ns in enumerate(f["nmod_moduli"]):
        n = int(ns)
        a, b, z = f["nmod_a%d" % k], f["nmod_b%d" % k], f["nmod_z%d" % k]
        zn = len(a) + len(b) - 1
        assert np.array_equal(O.nmod_poly_mul_mid(a, b, 0, zn, n), z)
        lo, hi = zn // 3, zn + 5
        mid = O.nmod_poly_mul_mid(a, b, lo, hi, n)
        assert np.array_equal(mid[: zn - lo], z[lo:])
        assert not mid[zn - lo:].any()


def test_fmpz_oracle_vs_fixtures():
    f = np.load(os.path.join(GOLDEN, "poly_small.npz"))
    for k, (aw, bw, zw) in enumerate(f["fmpz_widths"]):
        a, b, z = f["fmpz_a%d" % k], f["fmpz_b%d" % k], f["fmpz_z%d" % k]
        an, bn = len(a) // aw, len(b) // bw
        zn = an + bn - 1
        ok, out = O.fmpz_poly_mul_mid(a, int(aw), b, int(bw), 0, zn, int(zw))
        assert ok == 1
        assert np.array_equal(out, z)


def test_fmpz_oracle_declines_above_eight_primes():
    """fmpz_poly_mul.c:377-379: more than 8 primes -> return 0"""
    w = 4   # 200-bit coefficients: 2*200+1+log2(bn) > 397 bits
    an = bn = 10
    a = np.zeros(an * w, dtype=np.uint64)
    b = np.zeros(bn * w, dtype=np.uint64)
    a[w - 1::w] = np.uint64(1 << 7)
    b[w - 1::w] = np.uint64(1 << 7)
    assert O.fmpz_poly_nprimes(a, w, b, w) == 0
    ok, _ = O.fmpz_poly_mul_mid(a, w, b, w, 0, an + bn - 1, 8)
    assert ok == 0


def test_poly_ks_reference_small():
    """the Kronecker fixture generator against schoolbook"""
    rnd = random.Random(9)
    for _ in range(20):
        a = [rnd.randrange(-10**20, 10**20) for _ in range(rnd.randint(1, 9))]
        b = [rnd.randrange(-10**20, 10**20) for _ in range(rnd.randint(1, 9))]
        sb = [0] * (len(a) + len(b) - 1)
        for i, x in enumerate(a):
            for j, y in enumerate(b):
                sb[i + j] += x * y
        assert poly_ks(a, b, signed=True) == sb
