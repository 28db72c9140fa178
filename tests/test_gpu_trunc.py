"""GPU parity of the truncated transforms (sd_fft_trunc / sd_ifft_trunc at
row granularity, sd_ifft.c:1320-1493): a product of zn coefficients on a
2^depth-point transform (depth 13..23) runs the row pass over the first
ceil(zn / 4096) rows only and inverts the columns with the truncated inverse
transform.  Every case is checked against the oracle (or Python integers)
and against the same product with truncation switched off, over every
column depth d1 = 1..11 and row counts across (2^(d1-1), 2^d1]: the full
top block, and the boundary chains through both halves of every level."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_limbs, to_int

pytestmark = pytest.mark.gpu


def _lengths(d1):
    """(an, bn) pairs whose product has zn = an + bn - 1 coefficients with
    ceil(zn / 4096) spread over (2^(d1-1), 2^d1]"""
    lo, hi = (1 << (d1 - 1)) + 1, 1 << d1
    ts = sorted({lo, hi, (lo + hi) // 2, lo + 1 if lo < hi else lo})
    out = []
    for t in ts:
        zn = (t - 1) * 4096 + 1 + (t * 997) % 4096
        an = zn // 3 + 1
        out.append((an, zn + 1 - an))
    return out


@pytest.mark.parametrize("d1", range(1, 12))
def test_nmod_truncated_products(d1):
    import flint_amd as fa
    rng = np.random.default_rng(d1)
    n = 0x0003F00000000001 if d1 % 2 else (1 << 61) - 1
    for an, bn in _lengths(d1):
        a = rng.integers(0, n, an, dtype=np.uint64)
        b = rng.integers(0, n, bn, dtype=np.uint64)
        zn = an + bn - 1
        print("d1 %d an %d bn %d rows %d" % (d1, an, bn, -(-zn // 4096)), flush=True)
        fa.set_truncation(2)           # wherever it applies
        try:
            got = fa.nmod_poly_mul_mid(a, b, 0, zn, n)
            fa.set_truncation(0)
            full = fa.nmod_poly_mul_mid(a, b, 0, zn, n)
        finally:
            fa.set_truncation(1)
        assert np.array_equal(got, full), (an, bn)
        if zn <= 300000:
            assert np.array_equal(got, O.nmod_poly_mul_mid(a, b, 0, zn, n)), (an, bn)


@pytest.mark.parametrize("an,bn", [(3000, 2000), (9000, 300), (40001, 39999), (100000, 60000),
                                   (700000, 300001), (1 << 21, 1 << 20)])
def test_mpn_truncated_products(an, bn):
    import flint_amd as fa
    rng = np.random.default_rng(an)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    p = fa.mpn_plan(an, bn)
    print(an, bn, p, flush=True)
    fa.set_truncation(2)
    try:
        got = fa.mpn_mul(a, b)
        sq = fa.mpn_mul(a, a)
        fa.set_truncation(0)
        full = fa.mpn_mul(a, b)
    finally:
        fa.set_truncation(1)
    assert np.array_equal(got, full)
    if an + bn <= 300000:
        assert to_int(got) == to_int(a) * to_int(b)
    else:
        for q in ((1 << 61) - 1, 1152921504606846883):
            assert O.mpn_mod_q(got, q) == O.mpn_mod_q(a, q) * O.mpn_mod_q(b, q) % q
    # a square takes the one-transform route through the same truncation
    if an <= 100000:
        assert to_int(sq) == to_int(a) ** 2


def test_batch_truncated_matches_single_calls():
    """a batch mixing depths 13..19 (row lists cut per job) against single
    products with truncation off"""
    import flint_amd as fa
    rng = np.random.default_rng(5)
    reqs = []
    for an, bn, n in ((5000, 4000, (1 << 61) - 1), (20000, 9000, 65537), (70000, 70000, 1000003),
                      (3000, 130000, 0x0003F00000000001), (200000, 100000, (1 << 59) + 55)):
        while not O.lib().or_is_prime(n):
            n += 2
        a = rng.integers(0, n, an, dtype=np.uint64)
        b = rng.integers(0, n, bn, dtype=np.uint64)
        reqs.append({"kind": "nmod", "a": a, "b": b, "n": n, "zl": 0, "zh": an + bn - 1})
    for an, bn in ((6000, 5000), (90000, 30000)):
        reqs.append({"kind": "mpn", "a": rand_limbs(rng, an), "b": rand_limbs(rng, bn)})
    fa.set_truncation(2)                    # every job truncated where it applies
    try:
        res = fa.batch_mul(reqs)
        fa.set_truncation(0)
        for q, (st, z) in zip(reqs, res):
            assert st == 1
            if q["kind"] == "mpn":
                assert np.array_equal(z, fa.mpn_mul(q["a"], q["b"]))
            else:
                assert np.array_equal(z, fa.nmod_poly_mul_mid(q["a"], q["b"], 0, q["zh"], q["n"]))
    finally:
        fa.set_truncation(1)
