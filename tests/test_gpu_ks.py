"""GPU parity of the Kronecker-substitution route for fmpz products beyond
eight primes: exact Python-integer products on random signed inputs (mixed
signs, all positive, all negative, sparse, extreme magnitudes), windows,
agreement with the eight-prime path where that one computes, and agreement
of the signed-slot (one product) and non-negative-part (three products)
routes, with the borrow and sign corner cases of the former."""
import numpy as np
import pytest

from tests.golden.make_fixtures import poly_ks

pytestmark = pytest.mark.gpu


def enc(vals, w):
    return np.frombuffer(b"".join((x % (1 << (64 * w))).to_bytes(8 * w, "little") for x in vals),
                         dtype=np.uint64).copy()


def dec(limbs, w):
    out = []
    for i in range(len(limbs) // w):
        x = int.from_bytes(np.ascontiguousarray(limbs[i * w:(i + 1) * w]).tobytes(), "little")
        if x >> (64 * w - 1):
            x -= 1 << (64 * w)
        out.append(x)
    return out


def draw(rng, n, bits, sign):
    out = []
    for _ in range(n):
        x = int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits)
        if sign == "neg" or (sign == "mixed" and rng.integers(0, 2)):
            x = -x
        out.append(x)
    return out


CASES = [(1, 1, 64, "mixed"), (7, 5, 200, "mixed"), (100, 90, 300, "neg"), (300, 17, 600, "pos"),
         (1000, 800, 130, "mixed"), (2000, 1500, 1000, "mixed"), (64, 64, 4096, "mixed")]


@pytest.mark.parametrize("an,bn,bits,sign", CASES)
def test_ks_exact(an, bn, bits, sign):
    import flint_amd as fa
    rng = np.random.default_rng(an * 7 + bits)
    a = draw(rng, an, bits, sign)
    b = draw(rng, bn, bits, "mixed")
    w = bits // 64 + 1
    zw = (2 * bits + 20) // 64 + 2
    want = poly_ks(a, b, signed=True)
    zn = an + bn - 1
    z = fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, 0, zn + 3, zw)
    got = dec(z, zw)
    assert got[:zn] == want and got[zn:] == [0, 0, 0]
    lo, hi = zn // 3, zn // 3 + 11
    zwin = fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, lo, hi, zw)
    assert dec(zwin, zw) == (want + [0] * 20)[lo:hi]


def test_ks_extremes_and_zeros():
    import flint_amd as fa
    bits = 257
    top = (1 << bits) - 1
    a = [top, -top, 0, 0, top, -1, 1, 0]
    b = [-top, 0, top, -top]
    w = bits // 64 + 1
    zw = (2 * bits + 8) // 64 + 2
    want = poly_ks(a, b, signed=True)
    got = dec(fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, 0, len(a) + len(b) - 1, zw), zw)
    assert got == want


def test_ks_matches_eight_prime_path():
    import flint_amd as fa
    rng = np.random.default_rng(99)
    bits, an, bn = 120, 3000, 2500
    w = bits // 64 + 1
    zw = (2 * bits + 20) // 64 + 2
    a = enc(draw(rng, an, bits, "mixed"), w)
    b = enc(draw(rng, bn, bits, "mixed"), w)
    z8 = fa.fmpz_poly_mul_mid(a, w, b, w, 0, an + bn - 1, zw)
    assert z8 is not None
    zk = fa.fmpz_poly_mul_ks(a, w, b, w, 0, an + bn - 1, zw)
    assert np.array_equal(z8, zk)


SIGNED_CASES = [
    # (name, a, b): leading zero coefficients, a negative top, zero runs that
    # carry a slot borrow, an all-zero operand, single coefficients
    ("lead-zeros", [5, -3, 0, 7, 0, 0], [-(1 << 70), 2, 0]),
    ("neg-top", [1, 2, 3, -4], [-1, -1, -(1 << 90)]),
    ("zero-runs", [-1, 0, 0, 0, -1, 0, 0, 5], [3, 0, 0, -2]),
    ("all-zero", [0, 0, 0], [7, -9]),
    ("single", [-(1 << 100)], [-(1 << 100) + 1]),
    ("min-max", [(1 << 129) - 1, -((1 << 129) - 1), -1, 1], [-((1 << 129) - 1), 1, (1 << 129) - 1]),
]


@pytest.mark.parametrize("name,a,b", SIGNED_CASES, ids=[c[0] for c in SIGNED_CASES])
def test_ks_signed_slots_edge_cases(name, a, b):
    """the signed-slot route (one integer product) against exact integers
    and the three-product route, on borrow and sign corner cases"""
    import flint_amd as fa
    w = 3
    zw = 6
    zn = len(a) + len(b) - 1
    want = poly_ks(a, b, signed=True)
    try:
        out = {}
        for v in (0, 1):
            fa.set_ks_variant(v)
            out[v] = dec(fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, 0, zn + 2, zw), zw)
    finally:
        fa.set_ks_variant(1)
    assert out[1][:zn] == want and out[1][zn:] == [0, 0]
    assert out[0] == out[1]


@pytest.mark.parametrize("an,bn,bits", [(3000, 2000, 700), (500, 4000, 2500), (1, 5000, 300)])
def test_ks_variants_agree(an, bn, bits):
    """one-product and three-product routes give the same words, windows
    included, on random signed inputs with some zero coefficients"""
    import flint_amd as fa
    rng = np.random.default_rng(an + bn + bits)
    a = draw(rng, an, bits, "mixed")
    b = draw(rng, bn, bits, "mixed")
    for k in range(0, an, 7):
        a[k] = 0
    w = bits // 64 + 1
    zw = (2 * bits + 20) // 64 + 2
    zn = an + bn - 1
    lo, hi = zn // 5, zn // 5 + 300
    try:
        out = {}
        for v in (0, 1):
            fa.set_ks_variant(v)
            out[v] = (fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, 0, zn, zw),
                      fa.fmpz_poly_mul_ks(enc(a, w), w, enc(b, w), w, lo, hi, zw))
    finally:
        fa.set_ks_variant(1)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[1][1], out[1][0][lo * zw:hi * zw])
