"""GPU sd_fft_ctx surface: the invariants of t-sd_fft.c (evaluation at
sd_fft_ctx_w(Q, i) in trunc_index order; ifft(fft(x)) = 2^L x) and
agreement with the oracle's transforms, for the prime of t-sd_fft.c:100,
257 (t-sd_fft.c:107) and the product primes."""
import random

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def horner(X, n, pt, p):
    r = 0
    for c in reversed(X[:n]):
        r = (r * pt + int(c)) % p
    return r


@pytest.mark.parametrize("p,maxL", [(0x0003f00000000001, 16), (257, 8)])
def test_sd_fft_trunc_invariants(p, maxL):
    import flint_amd as fa
    Q = fa.SdFft(p)
    rnd = random.Random(p % 1000 + maxL)
    for L in range(0, maxL + 1):
        N = 1 << L
        for rep in range(3):
            X = [rnd.randrange(p) for _ in range(N)]
            it = N if rep == 0 else rnd.randint(1, N)
            ot = N if rep == 0 else rnd.randint(1, N)
            d = np.zeros(N)
            d[:it] = X[:it]
            out = Q.fft_trunc(d, L, it, ot)
            for _ in range(5):
                i = rnd.randrange(ot)
                pt = int(Q.w(i)) % p
                assert pt == O.lib().or_w(p, i)
                assert int(out[O.lib().or_trunc_index(L, i)]) % p == horner(X, it, pt, p)
            tr = N if rep == 0 else rnd.randint(1, N)
            d = np.zeros(N)
            d[:tr] = X[:tr]
            f = Q.fft_trunc(d, L, tr, tr)
            y = Q.ifft_trunc(f, L, tr)
            m = pow(2, L, p)
            for i in range(tr):
                assert int(y[i]) % p == (X[i] * m) % p, (L, tr, i)


@pytest.mark.parametrize("L", [12, 13, 18, 20])
def test_sd_fft_matches_oracle(L):
    import flint_amd as fa
    p = O.default_primes()[2]
    Q = fa.SdFft(p)
    rng = np.random.default_rng(L)
    N = 1 << L
    X = rng.integers(0, p, N, dtype=np.uint64)
    for it, ot in ((N, N), (N // 2 + 300, N - 1000), (N // 2 + 1, N // 2 + 1)):
        d = np.zeros(N, dtype=np.uint64)
        d[:it] = X[:it]
        ref = O.fft_trunc(p, d, L, it, ot)
        out = Q.fft_trunc(d.astype(np.float64), L, it, ot)
        ot_r = ((ot + 255) // 256) * 256
        assert np.array_equal(out[:ot_r].astype(np.uint64), ref[:ot_r])
        back = Q.ifft_trunc(out, L, ot)
        refb = O.ifft_trunc(p, ref, L, ot)
        # the inverse reads block-rounded truncations; compare the valid part
        assert np.array_equal(back[:ot_r].astype(np.uint64), refb[:ot_r])


@pytest.mark.parametrize("p", [0x0003f00000000001, 257])
def test_sd_point_mul_and_convolution(p):
    """sd_fft_ctx_point_mul / _point_sqr: a m b and a m a mod p pointwise;
    fft -> point_mul with m = 2^-L -> ifft is the product mod x^N - 1 of
    the inputs, against schoolbook arithmetic"""
    import flint_amd as fa
    Q = fa.SdFft(p)
    rnd = random.Random(p % 997)
    for L in (4, 8, 11):
        if p == 257 and L > 8:
            continue
        N = 1 << L
        a = [rnd.randrange(p) for _ in range(N)]
        b = [rnd.randrange(p) for _ in range(N)]
        m = rnd.randrange(1, p)
        got = Q.point_mul(np.array(a, dtype=np.float64), np.array(b, dtype=np.float64), m, L)
        assert [int(x) for x in got] == [x * m * y % p for x, y in zip(a, b)]
        got = Q.point_sqr(np.array(a, dtype=np.float64), m, L)
        assert [int(x) for x in got] == [x * m * x % p for x in a]
        A = Q.fft_trunc(np.array(a, dtype=np.float64), L, N, N)
        B = Q.fft_trunc(np.array(b, dtype=np.float64), L, N, N)
        C = Q.point_mul(A, B, pow(N, -1, p), L)
        c = Q.ifft_trunc(C, L, N)
        want = [0] * N
        for i in range(N):
            for j in range(N):
                want[(i + j) % N] = (want[(i + j) % N] + a[i] * b[j]) % p
        assert [int(x) % p for x in c] == want
