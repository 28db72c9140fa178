"""The oracle's integer windows (or_mpn_mul_range, the restatement of
_mpn_ctx_mpn_mul_range, mpn_mul.c:1617-1978 and mpn_mul_np2.c:247-433)
against the window computed from its definition with Python integers.

The definition is the reference's, worked out from the code: with chunk
coefficients c_j = sum_{i+k=j} A_i B_k (A_i = bits [i b, (i+1) b) of a,
b the plan's chunk width) and T = 2^(64 lo),
- two-prime plans (max(an, bn) <= 40000): W = sum_j floor(c_j 2^(j b) / T),
  i.e. (P - F) / T with F = sum over j b < 64 lo of (c_j 2^(j b) mod T);
- wider plans: W = floor(S / T) with S = sum_{j >= c_lo} c_j 2^(j b),
  c_lo = ceil((lo - n) 64 / b) for lo > n (n = limbs of M * np), else 0;
  S = P - (the low convolution sum_{i+k < c_lo} A_i B_k 2^((i+k) b));
both taken mod 2^(64 (hi - lo)).  The Python side never forms a residue,
a transform or a CRT: it shares nothing with the oracle but the plan
(np, bits), whose rule is pinned in test_oracle.py.  The checks of the
reference's own t-mulmid.c (the deficit exact - W is below
min(an, bn, lo) units of limb lo) run on every window, and a planted
wide-plan case makes the dropped carry actually occur.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import from_int, rand_limbs, to_int


def _coeff_len(np_):
    M = 1
    for p in O.default_primes()[:np_]:
        M *= int(p)
    return ((M * np_).bit_length() + 63) // 64


def _chunks(x, bits, n):
    mask = (1 << bits) - 1
    return [(x >> (i * bits)) & mask for i in range(n)]


def _low_conv(x, y, ylen, bits, c):
    """sum over i + k < c of X_i Y_k 2^((i+k) bits) where X, Y are the
    bit-chunks of x and y; loops over the chunks of y (the shorter one)"""
    s = 0
    for k in range(min(c, ylen)):
        yk = (y >> (k * bits)) & ((1 << bits) - 1)
        s += (yk << (k * bits)) * (x & ((1 << ((c - k) * bits)) - 1))
    return s


def window_definition(a, b, lo, hi, nthreads=1):
    an, bn = len(a), len(b)
    x, y = to_int(a), to_int(b)
    if bn > an:
        x, y, an, bn = y, x, bn, an
    pl = O.mpn_plan(len(a), len(b), nthreads)
    bits, np_ = pl["bits"], pl["np"]
    alen, blen = -(-64 * an // bits), -(-64 * bn // bits)
    zlen = alen + blen - 1
    w = hi - lo
    P = x * y
    if np_ == 2:
        E = min(zlen, -(-64 * lo // bits))
        mlen = min(alen, blen)
        guard = 2 * bits + mlen.bit_length()
        ka = 0 if 64 * lo < guard else min(E, (64 * lo - guard) // bits + 1)
        F = _low_conv(x, y, blen, bits, ka)
        A, B = _chunks(x, bits, alen), _chunks(y, bits, blen)
        T = 1 << (64 * lo)
        for j in range(ka, E):
            cj = sum(A[i] * B[j - i] for i in range(max(0, j - blen + 1), min(j, alen - 1) + 1))
            F += (cj << (j * bits)) % T
        assert (P - F) % T == 0
        W = (P - F) >> (64 * lo)
    else:
        nn = _coeff_len(np_)
        c_lo = 0 if lo <= nn else min(zlen, -(-(lo - nn) * 64 // bits))
        W = (P - _low_conv(x, y, blen, bits, c_lo)) >> (64 * lo)
    return W % (1 << (64 * w)), np_, bits


def _windows(rng, zn, k):
    out = {(0, zn), (zn - 1, zn), (0, 1)}
    for _ in range(k):
        lo = int(rng.integers(0, zn))
        out.add((lo, lo + 1 + int(rng.integers(0, zn - lo))))
    return sorted(out)


def _check(a, b, lo, hi, nthreads=1):
    got = O.mpn_mul_range(a, b, lo, hi, nthreads)
    want, np_, bits = window_definition(a, b, lo, hi, nthreads)
    assert to_int(got) == want, (len(a), len(b), lo, hi, np_, bits)
    # t-mulmid.c: 0 <= exact - W, below min(an, bn, lo) units of limb lo
    exact = (to_int(a) * to_int(b) >> (64 * lo)) % (1 << (64 * (hi - lo)))
    d = (exact - want) % (1 << (64 * (hi - lo)))
    assert d <= min(len(a), len(b), lo), (len(a), len(b), lo, hi, d)
    return d


def test_two_prime_windows_match_definition():
    rng = np.random.default_rng(71)
    shapes = [(int(rng.integers(1, 40)), int(rng.integers(1, 40))) for _ in range(25)]
    shapes += [(600, 17), (17, 600), (1500, 1499), (3000, 64)]
    nonzero = 0
    for an, bn in shapes:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        for lo, hi in _windows(rng, an + bn, 4):
            nonzero += _check(a, b, lo, hi) != 0
    # the two-prime window drops carries of the masked chunk tails often
    assert nonzero > 5


@pytest.mark.parametrize("an,bn", [(40001, 1), (50000, 3), (45000, 37), (70000, 200)])
def test_wide_plan_windows_match_definition(an, bn):
    rng = np.random.default_rng(an + bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    assert O.mpn_plan(an, bn)["np"] >= 4
    for lo, hi in _windows(rng, an + bn, 3) + [(an - 7, an + 2), (an + bn - 3, an + bn)]:
        _check(a, b, lo, hi)


def test_wide_plan_planted_dropped_carry():
    """b = 2^64 - 1 (one limb): the bits of a in [c_lo bits, 64 lo) are set
    so that the coefficients from c_lo up sum to T - 2^(c_lo bits) below T;
    the coefficients below c_lo then carry into limb lo, which the wide
    plan drops: its window is one below the exact one."""
    rng = np.random.default_rng(3)
    for an, lo, nthreads in ((50000, 30011, 1), (41000, 777, 1), (60000, 45000, 8)):
        a = rand_limbs(rng, an)
        b = np.array([2**64 - 1], dtype=np.uint64)
        pl = O.mpn_plan(an, 1, nthreads)
        bits, nn = pl["bits"], _coeff_len(pl["np"])
        c_lo = -(-(lo - nn) * 64 // bits)
        r = 64 * lo - c_lo * bits
        m = (-pow(2**64 - 1, -1, 1 << r)) % (1 << r)
        x = to_int(a)
        x &= ~(((1 << r) - 1) << (c_lo * bits))
        x |= m << (c_lo * bits)
        a = from_int(x, an)
        for hi in (lo + 1, lo + 40, an + 1):
            assert _check(a, b, lo, hi, nthreads) == 1, (an, lo, hi)


def test_full_windows_are_exact():
    rng = np.random.default_rng(5)
    for an, bn in ((5, 3), (900, 800), (41000, 20)):
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        assert to_int(O.mpn_mul_range(a, b, 0, an + bn)) == to_int(a) * to_int(b)
        nn = _coeff_len(O.mpn_plan(an, bn)["np"])
        if O.mpn_plan(an, bn)["np"] > 2:
            # lo <= n: c_lo = 0, the window is exact
            lo = min(nn, an + bn - 1)
            assert to_int(O.mpn_mul_range(a, b, lo, an + bn)) == (to_int(a) * to_int(b)) >> (64 * lo)
