"""GPU parity of polynomial division with remainder mod n
(_nmod_poly_divrem_mpn_ctx and the precomputed form, nmod_poly_mul.c:
1130-1313; the reference's t-nmod_poly_divrem.c checks a = b q + r).

Division by b with an invertible leading coefficient has one result, so the
checks pin it: a = b q + r with deg r < deg b, b q taken from the oracle's
product; schoolbook division on small cases; the inverse series satisfies
f g = 1 mod x^m."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _draw(rng, n, k):
    if n < 2**63:
        return rng.integers(0, n, k, dtype=np.uint64)
    return np.array([int.from_bytes(rng.bytes(8), "little") % n for _ in range(k)],
                    dtype=np.uint64)


def _schoolbook_divrem(a, b, n):
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    inv = pow(b[-1], -1, n)
    q = [0] * (len(a) - len(b) + 1)
    for i in range(len(q) - 1, -1, -1):
        c = a[i + len(b) - 1] * inv % n
        q[i] = c
        for j in range(len(b)):
            a[i + j] = (a[i + j] - c * b[j]) % n
    return q, a[:len(b) - 1]


def _check(a, b, q, r, n):
    bq = O.nmod_poly_mul_mid(b, q, 0, len(a), n)
    lhs = np.array([(int(x) + int(y)) % n for x, y in
                    zip(bq, np.concatenate([r, np.zeros(len(a) - len(r), dtype=np.uint64)]))],
                   dtype=np.uint64)
    return np.array_equal(lhs, a)


@pytest.mark.parametrize("n", [(1 << 61) - 1, 65537, 0x0003F00000000001, 2**64 - 59, 1 << 40])
@pytest.mark.parametrize("an,bn", [(10, 3), (300, 200), (5000, 1000), (40000, 39000),
                                   (100000, 30)])
def test_divrem_identity(n, an, bn):
    import flint_amd as fa
    rng = np.random.default_rng(an * 7 + bn)
    a, b = _draw(rng, n, an), _draw(rng, n, bn)
    if n % 2 == 0:
        b[-1] |= np.uint64(1)          # an odd leading coefficient is invertible mod 2^40
    q, r = fa.nmod_poly_divrem(a, b, n)
    assert len(q) == an - bn + 1 and len(r) == bn - 1
    assert int(q.max()) < n and (bn == 1 or int(r.max()) < n)
    assert _check(a, b, q, r, n)
    if an <= 300:
        sq, sr = _schoolbook_divrem(a, b, n)
        assert [int(x) for x in q] == sq and [int(x) for x in r] == sr


def test_inv_series_and_precomp():
    import flint_amd as fa
    n = (1 << 61) - 1
    rng = np.random.default_rng(5)
    f = _draw(rng, n, 3000)
    f[0] = np.uint64(12345)
    for m in (1, 2, 100, 4097, 20000):
        g = fa.nmod_poly_inv_series(f, m, n)
        fg = O.nmod_poly_mul_mid(f[:m], g, 0, m, n)
        assert int(fg[0]) == 1 and not fg[1:].any(), m
    b = _draw(rng, n, 700)
    M = fa.DivremPrecomp(b, 5000, n)
    for an in (700, 1000, 5699):
        a = _draw(rng, n, an)
        q, r = M.divrem(a)
        q2, r2 = fa.nmod_poly_divrem(a, b, n)
        assert np.array_equal(q, q2) and np.array_equal(r, r2)
        assert _check(a, b, q, r, n)
    assert M.divrem(_draw(rng, n, 5700)) is None      # quotient longer than the precision
    M.close()


def test_divrem_rejects_non_invertible_leading_coefficient():
    import flint_amd as fa
    a = np.arange(1, 20, dtype=np.uint64)
    b = np.array([1, 2, 6], dtype=np.uint64)          # 6 is not invertible mod 12
    with pytest.raises(fa.FlintAmdError):
        fa.nmod_poly_divrem(a, b, 12)


@pytest.mark.parametrize("n", [65537, (1 << 61) - 1, 2**64 - 59])
def test_mul_mid_precomp_windows(n):
    """_mul_precomp_init / _nmod_poly_mul_mid_precomp: b transformed once on
    2^depth points; windows of a b that fit the transform (zn <= N, or the
    wraparound zh <= N <= zn <= zl + N, with a longer than N folded) against
    the oracle's linear product, zero fill past zn, and 0 (nothing computed)
    for a window the transform cannot hold"""
    import flint_amd as fa
    rng = np.random.default_rng(n % 1000)

    def draw(m):
        return np.array([int.from_bytes(rng.bytes(8), "little") % n for _ in range(m)],
                        dtype=np.uint64)
    for depth, bn in ((4, 5), (9, 300), (13, 5000), (15, 20000)):
        N = 1 << depth
        b = draw(bn)
        M = fa.MulPrecomp(b, depth, n)
        an = N - bn + 1                       # zn = N: linear
        cases = [(an, 0, N, True), (an, N // 3, N // 3 + 7, True),
                 (an, 1, N + 5, True)]        # zh past zn: zero filled
        an = N - bn // 2 + 1                  # wraparound, zl = zn - N, zh = N
        cases.append((an, an + bn - 1 - N, N, True))
        an = N + (N - bn) // 2 + 1            # wraparound with a folded mod x^N - 1
        cases.append((an, an + bn - 1 - N, N, True))
        cases.append((2 * N + bn, 0, 10, False))     # does not fit
        for an, zl, zh, fits in cases:
            print(depth, bn, an, zl, zh, flush=True)
            a = draw(an)
            got = M.mul_mid(a, zl, zh)
            if not fits:
                assert got is None
                continue
            zn = an + bn - 1
            want = np.zeros(zh - zl, dtype=np.uint64)
            hi = min(zh, zn)
            if zl < hi:
                want[: hi - zl] = O.nmod_poly_mul_mid(a, b, zl, hi, n)
            assert got is not None and np.array_equal(got, want), (depth, an, zl, zh)
        M.close()
