"""GPU parity of the squaring route: when both operands are the same array
(mpn_mul.c:1665, nmod_poly_mul.c:939, fmpz_poly_mul.c:352) one forward
transform per prime is run and the pointwise stage squares
(sd_fft_ctx_point_sqr, mpn_mul.c:1425-1452).  Squares are checked against
the oracle and Python integers, against the general two-transform product
of the same operands, and at 2^24 limbs against residues mod 61-bit primes.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_limbs, to_int

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 300, 5000, 40000, 65536, 300000, 1 << 20, 3 << 20])
def test_mpn_square_matches_general_product(n):
    import flint_amd as fa
    rng = np.random.default_rng(n)
    a = rand_limbs(rng, n)
    b = a.copy()                       # same values, another array: general route
    print(n, fa.mpn_plan(n, n), flush=True)
    sq = fa.mpn_mul(a, a)
    gen = fa.mpn_mul(a, b)
    assert np.array_equal(sq, gen)
    if n <= 65536:
        assert np.array_equal(sq, O.mpn_mul(a, a))
    fa.set_square_route(False)
    try:
        assert np.array_equal(fa.mpn_mul(a, a), gen)
    finally:
        fa.set_square_route(True)


def test_mpn_square_all_ones_and_device_pointers():
    import torch
    import flint_amd as fa
    dev = torch.device("cuda:0")
    for n in (4096, 70000, 1 << 21):
        a = np.full(n, np.uint64(2**64 - 1))
        N = 64 * n
        want = (1 << (2 * N)) - (1 << (N + 1)) + 1       # (2^N - 1)^2
        got = fa.mpn_mul(a, a)
        assert to_int(got) == want, n
        # the device entry point with one pointer for both operands
        da = torch.from_numpy(a.view(np.int64)).to(dev)
        dz = torch.empty(2 * n, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()   # the upload runs on torch's stream, the product on s
        s = torch.cuda.Stream(device=dev)
        fa.mpn_mul_device(dz.data_ptr(), da.data_ptr(), n, da.data_ptr(), n, s.cuda_stream)
        s.synchronize()
        assert np.array_equal(dz.cpu().numpy().view(np.uint64), got), n


def test_mpn_square_2p24_mod_q():
    import flint_amd as fa
    n = 1 << 24
    a = fa.synth_limbs(0x5A5A, n)
    z = fa.mpn_mul(a, a)
    for q in ((1 << 61) - 1, 1152921504606846883, 4611686018427387847):
        r = O.mpn_mod_q(a, q)
        assert O.mpn_mod_q(z, q) == r * r % q, q


def test_nmod_and_fmpz_squares():
    import flint_amd as fa
    rng = np.random.default_rng(77)
    for n, ln in (((1 << 61) - 1, 3000), (0x0003F00000000001, 20000), (65537, 70000),
                  (2**64 - 59, 5000)):
        a = rng.integers(0, n, ln, dtype=np.uint64) if n < 2**63 else \
            (rng.integers(0, 2**63, ln, dtype=np.uint64) * np.uint64(2)) % np.uint64(n)
        zn = 2 * ln - 1
        sq = fa.nmod_poly_mul_mid(a, a, 0, zn, n)
        assert np.array_equal(sq, fa.nmod_poly_mul_mid(a, a.copy(), 0, zn, n)), n
        if ln <= 5000:
            assert np.array_equal(sq, O.nmod_poly_mul_mid(a, a, 0, zn, n)), n
    w = 3
    for ln in (50, 4000):
        vals = [int(v) * (1 << 70) + int(u) for v, u in
                zip(rng.integers(-2**62, 2**62, ln), rng.integers(0, 2**62, ln))]
        a = np.zeros(ln * w, dtype=np.uint64)
        for i, v in enumerate(vals):
            a[i * w:(i + 1) * w] = np.frombuffer((v % (1 << 192)).to_bytes(24, "little"),
                                                 dtype=np.uint64)
        zw = 5
        sq = fa.fmpz_poly_mul_mid(a, w, a, w, 0, 2 * ln - 1, zw)
        gen = fa.fmpz_poly_mul_mid(a, w, a.copy(), w, 0, 2 * ln - 1, zw)
        assert sq is not None and np.array_equal(sq, gen)
        if ln == 50:
            for k in range(2 * ln - 1):
                want = sum(vals[i] * vals[k - i] for i in range(max(0, k - ln + 1), min(k, ln - 1) + 1))
                v = to_int(sq[k * zw:(k + 1) * zw])
                if v >> (64 * zw - 1):
                    v -= 1 << (64 * zw)
                assert v == want, k
