"""CPU model of the nmod CRT of flint_amd/csrc/fa_poly.hip (fa_crt_nmod_one).

The kernel finds the mixed-radix digits t_k of the coefficient c in [0, M)
with FP64 mulmods (h = a b, q = rint(h / p), r = fma(-q, p, h) + fma(a, b, -h))
and reductions to [-p/2, p/2] and [0, p), then c mod n by Horner's rule with
128-bit products.  This model replays those double operations exactly (FMA
through rationals, rint half-even), asserts the operand bound |a b| < 2 p^2
at every mulmod and the high word below n at every 128-bit reduction, and
checks the digits against c itself and the result against c mod n, for
every prime count 1..8 at every offset into the eight default primes, with
c at 0, near M - 1 and random, stored residues in [0, p) and in the signed
range, and moduli n from 1 to 2^64 - 59.
"""
import random
from fractions import Fraction

import pytest

from oracle import oracle as O


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _mulmod(a, b, p, pinv):
    assert abs(Fraction(a) * Fraction(b)) < 2 * p * p
    h = a * b
    q = float(round(h * pinv))
    return _fma(-q, p, h) + _fma(a, b, -h)


def _pm1n(a, p, pinv):
    return _fma(-float(round(a * pinv)), p, a)


def _r0n(a, p, pinv):
    r = _pm1n(a, p, pinv)
    if r < 0:
        r += p
    assert 0 <= r < p and r == int(r)
    return float(r)


def _consts(primes):
    M = 1
    for p in primes:
        M *= p
    gcm = [(M // p) % p for p in primes]
    ginv, gpm = [], []
    for k, pk in enumerate(primes):
        pre, row = 1 % pk, []
        for j in range(k):
            row.append(primes[j] % pk)
            pre = pre * (primes[j] % pk) % pk
        gpm.append(row)
        ginv.append(pow(pre, pk - 2, pk))
    return M, gcm, ginv, gpm


def garner_model(xs, primes, n):
    """returns (c from the digits, the kernel's c mod n)"""
    M, gcm, ginv, gpm = _consts(primes)
    t = []
    for k, pk in enumerate(primes):
        p = float(pk)
        pinv = 1.0 / p
        x = _r0n(float(xs[k]), p, pinv)
        r = _mulmod(x, float(gcm[k]), p, pinv)
        if k == 0:
            t.append(_r0n(r, p, pinv))
            continue
        u = t[k - 1]
        for i in range(k - 2, -1, -1):
            u = _pm1n(_mulmod(u, float(gpm[k][i]), p, pinv) + t[i], p, pinv)
        u = _pm1n(u, p, pinv)
        t.append(_r0n(_mulmod(r - u, float(ginv[k]), p, pinv), p, pinv))
    c = 0
    for k in range(len(primes) - 1, -1, -1):
        c = c * primes[k] + int(t[k])
    v = int(t[-1]) % n
    for k in range(len(primes) - 2, -1, -1):
        w = v * (primes[k] % n) + int(t[k])
        assert (w >> 64) < n
        v = w % n
    return c, v


@pytest.mark.parametrize("np_", range(1, 9))
def test_garner_model_exact(np_):
    P = [int(p) for p in O.default_primes()]
    rnd = random.Random(np_)
    for off in range(0, 9 - np_):
        primes = P[off:off + np_]
        M = 1
        for p in primes:
            M *= p
        for trial in range(60):
            kind = trial % 4
            if kind == 0:
                c = rnd.randrange(M)
            elif kind == 1:
                c = M - 1 - rnd.randrange(1000)
            elif kind == 2:
                c = rnd.randrange(1000)
            else:
                c = rnd.randrange(1 << rnd.randint(1, M.bit_length() - 1))
            xs = []
            for p in primes:
                x = c * pow((M // p) % p, p - 2, p) % p
                if rnd.random() < 0.5 and x > p // 2:
                    x -= p
                xs.append(x)
            n = rnd.choice([1, 2, 3, (1 << 64) - 59, (1 << 63) + 29,
                            rnd.randrange(2, 1 << 64), rnd.randrange(2, 1 << 20)])
            got_c, got = garner_model(xs, primes, n)
            assert got_c == c
            assert got == c % n
