"""GPU parity of the small-modulus packed route (nmod_poly/mullow_fft_small.c:
24-429): for moduli n <= 23 whose operands pass the norm bounds, pairs of
coefficients are multiplied at half length with one FFT prime; the output
equals the oracle's product mod n on full products and windows (even and
odd window ends), squares take the one-operand route, and operands beyond
the bounds are declined (0) so the caller uses the plain route."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 3, 7, 16, 23])
def test_repack_matches_oracle(n):
    import flint_amd as fa
    rng = np.random.default_rng(n)
    taken = 0
    for an, bn in ((1, 1), (2, 3), (17, 40), (1000, 999), (3000, 2500), (30000, 20000)):
        a = rng.integers(0, n, an, dtype=np.uint64)
        b = rng.integers(0, n, bn, dtype=np.uint64)
        zn = an + bn - 1
        for lo, hi in ((0, zn), (1, zn), (zn // 3, zn - 1), (zn // 2 + 1, zn)):
            if lo >= hi:
                continue
            want = O.nmod_poly_mul_mid(a, b, lo, hi, n)
            st, z = fa.nmod_poly_mulmid_repack(a, b, lo, hi, n)
            if st == 1:
                taken += 1
                assert np.array_equal(z, want), (an, bn, lo, hi)
            assert np.array_equal(fa.nmod_poly_mulmid_fft_small(a, b, lo, hi, n), want)
        sq = fa.nmod_poly_mulmid_fft_small(a, a, 0, 2 * an - 1, n)
        assert np.array_equal(sq, O.nmod_poly_mul_mid(a, a, 0, 2 * an - 1, n))
    assert taken > 0


def test_repack_declines_outside_its_bounds():
    import flint_amd as fa
    rng = np.random.default_rng(1)
    a = rng.integers(0, 23, 300000, dtype=np.uint64)     # beyond the length table for 23
    assert fa.nmod_poly_mulmid_repack(a, a, 0, 10, 23)[0] == 0
    b = rng.integers(0, 1000, 50, dtype=np.uint64)
    assert fa.nmod_poly_mulmid_repack(b, b, 0, 99, 1009)[0] == 0   # modulus above 23
