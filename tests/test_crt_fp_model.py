"""CPU model of the radix-2^24 FP64 CRT of flint_amd/csrc/fa_crt_fp.h.

The GPU kernel forms sum_i x_i (M/p_i) - k M with exact FP64 FMAs on 24-bit
digits.  This model replays the same steps on Python integers and asserts,
at every FMA and carry step, that the value is an integer below 2^53 (so
the double operation is exact), then checks the recombined coefficient
against the residues' defining integer for every prime count 2..8, with the
residues in [0, p) and in [-p/2, p/2] (the kernel's),
including the cases where the quotient estimate rounds up (c near M/2 and
M - 1) and c = 0.
"""
import random
from fractions import Fraction

import pytest

from oracle import oracle as O

B = 1 << 24


def _exact(v):
    assert v.denominator == 1 and abs(v) < 2 ** 53, v
    return float(v)


def _digits(x, nd):
    return [float((x >> (24 * d)) & 0xFFFFFF) for d in range(nd)]


def crt_fp_model(xs, primes):
    np_ = len(primes)
    M = 1
    for p in primes:
        M *= p
    ND = (50 * np_ + 23) // 24
    NDC = (50 * (np_ - 1) + 23) // 24
    cof = [_digits(M // p, 16) for p in primes]
    Md = _digits(M, 18)
    assert all(c == 0 for row in cof for c in row[NDC:]) and all(c == 0 for c in Md[ND:])
    rp = [1.0 / p for p in primes]
    col = [0.0] * ND

    def carry():
        c = 0.0
        for m in range(ND):
            v = _exact(Fraction(col[m]) + Fraction(c))
            c = float((Fraction(v) / B).__floor__())
            col[m] = _exact(Fraction(v) - Fraction(c) * B)
        return c

    qe = 0.0
    for i, x in enumerate(xs):
        if i == 6:
            top = carry()
            col[ND - 1] = _exact(Fraction(col[ND - 1]) + Fraction(top) * B)
        x = float(x)
        qe = float(Fraction(x) * Fraction(rp[i]) + Fraction(qe))   # one rounding (fma)
        v = float(int(x) >> 24)
        u = _exact(Fraction(x) - Fraction(v) * B)
        for d in range(NDC):
            col[d] = _exact(Fraction(u) * Fraction(cof[i][d]) + Fraction(col[d]))
            col[d + 1] = _exact(Fraction(v) * Fraction(cof[i][d]) + Fraction(col[d + 1]))
    k = float(round(qe))          # rint, ties to even
    for m in range(ND):
        col[m] = _exact(Fraction(col[m]) - Fraction(k) * Fraction(Md[m]))
    if carry() < 0:
        for m in range(ND):
            col[m] += Md[m]
        carry()
    return sum(int(col[m]) << (24 * m) for m in range(ND))


@pytest.mark.parametrize("np_", range(2, 9))
def test_crt_fp_model_exact(np_):
    primes = [int(p) for p in O.default_primes()[:np_]]
    M = 1
    for p in primes:
        M *= p
    cofinv = [pow(M // p % p, -1, p) for p in primes]
    rs = random.Random(np_)
    cases = [0, 1, M - 1, M // 2, M // 2 - 1, M // 2 + 1, M // 3]
    cases += [rs.randrange(M) for _ in range(60)]
    cases += [rs.randrange(2 ** 40) for _ in range(10)]
    cases += [M - 1 - rs.randrange(2 ** 40) for _ in range(10)]
    for c in cases:
        xs = [c * cofinv[i] % primes[i] for i in range(np_)]
        assert crt_fp_model(xs, primes) == c
        # the kernel's representatives in [-p/2, p/2]
        xs = [x - p if 2 * x > p else x for x, p in zip(xs, primes)]
        assert crt_fp_model(xs, primes) == c
