"""GPU parity of _nmod_poly_mul_mid / _fmpz_poly_mul_mid (drop-ins)
against the golden fixtures (Kronecker products on Python integers) and the
oracle, including windows, zero padding past the product and the decline
contract of the fmpz entry (fmpz_poly_mul.c:377-379)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.make_fixtures import poly_ks

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_nmod_vs_fixtures_and_windows():
    import flint_amd as fa
    f = np.load(os.path.join(GOLDEN, "poly_small.npz"))
    for k, ns in enumerate(f["nmod_moduli"]):
        n = int(ns)
        a, b, z = f["nmod_a%d" % k], f["nmod_b%d" % k], f["nmod_z%d" % k]
        zn = len(a) + len(b) - 1
        print("nmod", n, len(a), len(b), flush=True)
        assert np.array_equal(fa.nmod_poly_mul_mid(a, b, 0, zn, n), z)
        lo, hi = zn // 3, zn + 7
        mid = fa.nmod_poly_mul_mid(a, b, lo, hi, n)
        assert np.array_equal(mid[: zn - lo], z[lo:])
        assert not mid[zn - lo:].any()


@pytest.mark.parametrize("bits", [1, 2, 20, 31, 32, 33, 49, 50, 51, 52, 63, 64])
def test_nmod_random_moduli(bits):
    import flint_amd as fa
    rng = np.random.default_rng(bits)
    for rep in range(3):
        n = int(rng.integers(1 << (bits - 1), (1 << bits) - 1, dtype=np.uint64, endpoint=True))
        n = max(n, 1)
        an, bn = int(rng.integers(1, 5000)), int(rng.integers(1, 5000))
        a = (rng.integers(0, 2**63, an, dtype=np.uint64) % np.uint64(n)).astype(np.uint64)
        b = (rng.integers(0, 2**63, bn, dtype=np.uint64) % np.uint64(n)).astype(np.uint64)
        zn = an + bn - 1
        z = fa.nmod_poly_mul_mid(a, b, 0, zn, n)
        assert np.array_equal(z, O.nmod_poly_mul_mid(a, b, 0, zn, n))


def test_nmod_fft_prime_moduli_and_long():
    import flint_amd as fa
    rng = np.random.default_rng(5)
    for n in O.default_primes()[:3] + [7340033, 2**61 - 1]:
        an, bn = 70000, 65000
        a = (rng.integers(0, 2**63, an, dtype=np.uint64) % np.uint64(n)).astype(np.uint64)
        b = (rng.integers(0, 2**63, bn, dtype=np.uint64) % np.uint64(n)).astype(np.uint64)
        zn = an + bn - 1
        z = fa.nmod_poly_mul_mid(a, b, 1000, zn - 1000, n)
        ref = O.nmod_poly_mul_mid(a, b, 1000, zn - 1000, n)
        assert np.array_equal(z, ref)


def test_nmod_all_max_coefficients():
    import flint_amd as fa
    n = 2**64 - 59
    a = np.full(3000, np.uint64(n - 1))
    b = np.full(2000, np.uint64(n - 1))
    zn = 4999
    ref = [c % n for c in poly_ks([n - 1] * 3000, [n - 1] * 2000)]
    assert [int(x) for x in fa.nmod_poly_mul_mid(a, b, 0, zn, n)] == ref


def test_fmpz_vs_fixtures():
    import flint_amd as fa
    f = np.load(os.path.join(GOLDEN, "poly_small.npz"))
    for k, (aw, bw, zw) in enumerate(f["fmpz_widths"]):
        a, b, z = f["fmpz_a%d" % k], f["fmpz_b%d" % k], f["fmpz_z%d" % k]
        an, bn = len(a) // aw, len(b) // bw
        zn = an + bn - 1
        print("fmpz", k, aw, an, bn, flush=True)
        out = fa.fmpz_poly_mul_mid(a, int(aw), b, int(bw), 0, zn, int(zw))
        assert out is not None
        assert np.array_equal(out, z)


def test_fmpz_declines_like_the_reference():
    import flint_amd as fa
    w = 4
    an = bn = 10
    a = np.zeros(an * w, dtype=np.uint64)
    b = np.zeros(bn * w, dtype=np.uint64)
    a[w - 1::w] = np.uint64(1 << 7)
    b[w - 1::w] = np.uint64(1 << 7)
    zw = 8
    z = np.full((an + bn - 1) * zw, np.uint64(7))
    assert fa.fmpz_poly_mul_mid(a, w, b, w, 0, an + bn - 1, zw) is None
    ok, _ = O.fmpz_poly_mul_mid(a, w, b, w, 0, an + bn - 1, zw)
    assert ok == 0


@pytest.mark.parametrize("bits,signed", [(30, True), (62, False), (120, True), (180, True)])
def test_fmpz_random_vs_oracle(bits, signed):
    import flint_amd as fa
    rng = np.random.default_rng(bits)
    an, bn = 3000, 2500
    w = bits // 64 + 1
    zw = (2 * bits + 16) // 64 + 1

    def draw(n):
        v = []
        for _ in range(n):
            x = int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits)
            v.append(-x if signed and rng.integers(0, 2) else x)
        return v, np.frombuffer(b"".join((x % (1 << (64 * w))).to_bytes(8 * w, "little")
                                         for x in v), dtype=np.uint64)
    _, a = draw(an)
    _, b = draw(bn)
    zn = an + bn - 1
    out = fa.fmpz_poly_mul_mid(a, w, b, w, 5, zn + 3, zw)
    ok, ref = O.fmpz_poly_mul_mid(a, w, b, w, 5, zn + 3, zw)
    assert ok == 1 and out is not None
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("an,bn,cap,zl,zh", [(3000, 2500, 2000, 0, 5499), (5000, 40, 1000, 7, 3000),
                                             (1200, 1300, 700, 600, 2400)])
def test_nmod_blocked_vs_oracle(an, bn, cap, zl, zh):
    """nmod products in block products (forced by a small cap) whose window
    shares are added mod n on the device, against the oracle"""
    import flint_amd as fa
    rng = np.random.default_rng(an + cap)
    n = (1 << 61) - 1
    a = rng.integers(0, n, an, dtype=np.uint64)
    b = rng.integers(0, n, bn, dtype=np.uint64)
    a[::5] = n - 1
    ref = O.nmod_poly_mul_mid(a, b, zl, zh, n)
    try:
        fa.set_poly_block_cap(cap)
        out = fa.nmod_poly_mul_mid(a, b, zl, zh, n)
    finally:
        fa.set_poly_block_cap(0)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("bits,cap", [(62, 2000), (180, 1500)])
def test_fmpz_blocked_vs_oracle(bits, cap):
    """fmpz products in block products whose shares are added as two's
    complement coefficients on the device, against the oracle"""
    import flint_amd as fa
    rng = np.random.default_rng(bits + cap)
    an, bn = 3000, 2200
    w = bits // 64 + 1
    zw = (2 * bits + 16) // 64 + 1

    def draw(n):
        v = []
        for _ in range(n):
            x = int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits)
            v.append(-x if rng.integers(0, 2) else x)
        return np.frombuffer(b"".join((x % (1 << (64 * w))).to_bytes(8 * w, "little")
                                      for x in v), dtype=np.uint64)
    a, b = draw(an), draw(bn)
    zn = an + bn - 1
    ok, ref = O.fmpz_poly_mul_mid(a, w, b, w, 3, zn + 2, zw)
    try:
        fa.set_poly_block_cap(cap)
        out = fa.fmpz_poly_mul_mid(a, w, b, w, 3, zn + 2, zw)
    finally:
        fa.set_poly_block_cap(0)
    assert ok == 1 and out is not None
    assert np.array_equal(out, ref)


def test_nmod_beyond_one_transform():
    """a product of more than 2^24 output coefficients (no single transform
    holds it) against the sum of two products that fit, formed on the host:
    a b = a_lo b + x^h a_hi b"""
    import flint_amd as fa
    rng = np.random.default_rng(77)
    n = (1 << 60) + 33   # any modulus works; the sum below stays below 2^64
    an, bn = (1 << 23) + 5000, (1 << 23) + 3000
    a = rng.integers(0, n, an, dtype=np.uint64)
    b = rng.integers(0, n, bn, dtype=np.uint64)
    zn = an + bn - 1
    print("beyond-transform nmod", an, bn, flush=True)
    z = fa.nmod_poly_mul_mid(a, b, 0, zn, n)
    h = an // 2
    lo = fa.nmod_poly_mul_mid(a[:h], b, 0, h + bn - 1, n)
    hi = fa.nmod_poly_mul_mid(a[h:], b, 0, an - h + bn - 1, n)
    want = np.zeros(zn, dtype=np.uint64)
    want[:h + bn - 1] = lo
    s = want[h:] + hi
    want[h:] = np.where(s >= n, s - np.uint64(n), s)
    assert np.array_equal(z, want)


def _fft_prime(bits, v, rng):
    """a prime k 2^v + 1 of the given bit length"""
    while True:
        k = int(rng.integers(1 << (bits - v - 1), 1 << (bits - v)))
        n = k * (1 << v) + 1
        if n.bit_length() == bits and O.lib().or_is_prime(n):
            return n


def test_nmod_direct_fft_moduli():
    """plan.c:319-349: an FFT prime modulus whose 2-valuation covers the
    transform is transformed directly (one prime, no CRT).  Products and
    windows (wraparound included) for such moduli of 23 to 50 bits against
    the oracle, a 2^21-coefficient product by evaluation at random points,
    the device entry, and the counter that shows which path ran; moduli
    that are not FFT primes keep the CRT path"""
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(41)
    moduli = [998244353, 469762049, 7340033, _fft_prime(40, 24, rng), _fft_prime(50, 22, rng)]

    def draw(n, k):
        return np.array([int.from_bytes(rng.bytes(8), "little") % n for _ in range(k)],
                        dtype=np.uint64)
    for n in moduli:
        for an, bn, zl, zh in ((3000, 2500, 0, 5499), (40000, 30000, 17, 69999),
                               (70000, 3000, 20000, 72999), (300000, 200000, 123456, 123999)):
            a, b = draw(n, an), draw(n, bn)
            before = fa.lib().fa_nmod_direct_count()
            print(n, an, bn, zl, zh, flush=True)
            z = fa.nmod_poly_mul_mid(a, b, zl, zh, n)
            assert fa.lib().fa_nmod_direct_count() == before + 1, n
            assert np.array_equal(z, O.nmod_poly_mul_mid(a, b, zl, zh, n)), (n, an, bn)
    # a long product (depth 21) checked by evaluation
    n = 998244353
    an, bn = 1 << 20, (1 << 20) - 5
    a, b = draw(n, an), draw(n, bn)
    z = fa.nmod_poly_mul_mid(a, b, 0, an + bn - 1, n)
    for x in (3, 123456789, int(rng.integers(2, n))):
        assert O.nmod_poly_eval(z, x, n) == O.nmod_poly_eval(a, x, n) * O.nmod_poly_eval(b, x, n) % n
    # the device entry on the caller's stream
    a, b = draw(n, 5000), draw(n, 4000)
    da = torch.from_numpy(a.view(np.int64)).to("cuda:0")
    db = torch.from_numpy(b.view(np.int64)).to("cuda:0")
    dz = torch.zeros(8999, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    before = fa.lib().fa_nmod_direct_count()
    fa.nmod_poly_mul_mid_device(dz.data_ptr(), 0, 8999, da.data_ptr(), 5000, db.data_ptr(), 4000,
                                n, s.cuda_stream)
    s.synchronize()
    assert fa.lib().fa_nmod_direct_count() == before + 1
    assert np.array_equal(dz.cpu().numpy().view(np.uint64), O.nmod_poly_mul_mid(a, b, 0, 8999, n))
    # not FFT primes (or too small a 2-valuation): the CRT path
    for n in ((1 << 61) - 1, 1000003, 3 * (1 << 10) + 1):
        a, b = draw(n, 3000), draw(n, 2000)
        before = fa.lib().fa_nmod_direct_count()
        z = fa.nmod_poly_mul_mid(a, b, 0, 4999, n)
        assert fa.lib().fa_nmod_direct_count() == before
        assert np.array_equal(z, O.nmod_poly_mul_mid(a, b, 0, 4999, n))
