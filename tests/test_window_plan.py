"""CPU checks of the integer-window correction (fa_window.hip) without a GPU.

- The library's restatement of FLINT's plan rule (mpn_ctx_best_profile,
  _fft_small_np2_choose, crt_data_find_bn_bound / find_bits) gives the
  oracle's plan on every shape and thread count tried, and the boundary
  coefficients it asks for start at c_lo (wide plans).
- A Python model of the device correction, limb for limb as fa_k_win_delta
  computes it (the boundary coefficients summed from chunk products, limbs
  [L, lo) of the exact product, the comparison or the floor), turns the
  exact window into the oracle's window (or_mpn_mul_range) on two-prime and
  wide plans, the planted dropped carry included.
"""
import numpy as np
import pytest

import flint_amd as fa
from oracle import oracle as O
from tests.helpers import from_int, rand_limbs, to_int
from tests.test_oracle_window import _coeff_len


@pytest.mark.parametrize("nthreads", [1, 2, 3, 8])
def test_window_plan_matches_oracle_plan(nthreads):
    rng = np.random.default_rng(nthreads)
    shapes = [(int(rng.integers(1, 40000)), int(rng.integers(1, 40000))) for _ in range(20)]
    shapes += [(int(x), int(y)) for x, y in rng.integers(40001, 1 << 22, (30, 2))]
    shapes += [(1 << 24, 1 << 24), (40001, 1), (1 << 22, 3), (1 << 22, 1 << 21), (40000, 40000),
               (3 << 20, 700000), (1 << 26, 1 << 20)]
    for an, bn in shapes:
        lo = (an + bn) // 2
        two, bits, k0, nk = fa.mpn_window_plan(an, bn, lo, nthreads)
        pl = O.mpn_plan(an, bn, nthreads)
        assert (two, bits) == (pl["np"] == 2, pl["bits"]), (an, bn, nthreads)
        if not two:
            nn = _coeff_len(pl["np"])
            zlen = pl["alen"] + pl["blen"] - 1
            c_lo = min(zlen, -(-(lo - nn) * 64 // bits)) if lo > nn else 0
            E = min(zlen, -(-64 * lo // bits))
            assert (k0, nk) == ((c_lo, E - c_lo) if c_lo < E else (0, 0)), (an, bn)
    # nthreads = 0 and lo = 0: no correction
    assert fa.mpn_window_plan(5000, 4000, 3000, 0)[3] == 0
    assert fa.mpn_window_plan(50000, 4000, 0, 1)[3] == 0


def _model_window(a, b, lo, hi, nthreads=1, full=None):
    """the device correction on Python integers, step for step; full: the
    exact product's limbs when already known (else formed here)"""
    an, bn = len(a), len(b)
    if full is None:
        full = from_int(to_int(a) * to_int(b), an + bn)
    w = hi - lo
    exact = to_int(full[lo:hi])
    two, bits, k0, nk = fa.mpn_window_plan(an, bn, lo, nthreads)
    if nk == 0:
        return exact
    alen, blen = -(-64 * an // bits), -(-64 * bn // bits)
    mask = (1 << bits) - 1
    xb = np.asarray(a, dtype=np.uint64).tobytes() + bytes(32)
    yb = np.asarray(b, dtype=np.uint64).tobytes() + bytes(32)
    nb = bits // 8 + 2

    def chunk(buf, i):
        pos = i * bits
        return (int.from_bytes(buf[pos // 8:pos // 8 + nb], "little") >> (pos % 8)) & mask

    coef = []
    for k in range(k0, k0 + nk):
        s = 0
        for i in range(max(0, k - blen + 1), min(k, alen - 1) + 1):
            s += chunk(xb, i) * chunk(yb, k - i)
        assert s < 1 << 512
        coef.append(s)
    L = k0 * bits // 64
    m = lo - L
    Mm = 1 << (64 * m)
    Y = sum(c << ((k0 + j) * bits - 64 * L) for j, c in enumerate(coef)) % Mm
    Pmid = to_int(full[L:lo])
    if not two:
        d = int(Pmid < Y)
    else:
        R = sum((c << ((k0 + j) * bits - 64 * L)) % Mm for j, c in enumerate(coef))
        d = (((Pmid - Y) % Mm) + R) >> (64 * m)
    return (exact - d) % (1 << (64 * w))


def test_model_correction_matches_oracle_two_prime():
    rng = np.random.default_rng(12)
    shapes = [(int(rng.integers(1, 60)), int(rng.integers(1, 60))) for _ in range(30)]
    shapes += [(700, 300), (2000, 1999), (39999, 5)]
    hits = 0
    for an, bn in shapes:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        zn = an + bn
        for _ in range(4):
            lo = int(rng.integers(0, zn))
            hi = lo + 1 + int(rng.integers(0, zn - lo))
            want = to_int(O.mpn_mul_range(a, b, lo, hi))
            got = _model_window(a, b, lo, hi)
            assert got == want, (an, bn, lo, hi)
            hits += want != (to_int(a) * to_int(b) >> (64 * lo)) % (1 << (64 * (hi - lo)))
    assert hits > 5


def test_model_correction_matches_oracle_wide():
    rng = np.random.default_rng(13)
    for an, bn, nt in ((40001, 2, 1), (50000, 9, 1), (60000, 33, 8), (45000, 1, 3)):
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        for lo, hi in ((1, 3), (an // 2, an // 2 + 9), (an - 3, an + bn), (an + bn - 1, an + bn)):
            assert _model_window(a, b, lo, hi, nt) == to_int(O.mpn_mul_range(a, b, lo, hi, nt))
    # the planted dropped carry of test_oracle_window
    an, lo = 50000, 30011
    a = rand_limbs(rng, an)
    b = np.array([2**64 - 1], dtype=np.uint64)
    pl = O.mpn_plan(an, 1)
    bits, nn = pl["bits"], _coeff_len(pl["np"])
    c_lo = -(-(lo - nn) * 64 // bits)
    r = 64 * lo - c_lo * bits
    mm = (-pow(2**64 - 1, -1, 1 << r)) % (1 << r)
    xv = (to_int(a) & ~(((1 << r) - 1) << (c_lo * bits))) | (mm << (c_lo * bits))
    a = from_int(xv, an)
    want = to_int(O.mpn_mul_range(a, b, lo, lo + 5))
    assert _model_window(a, b, lo, lo + 5) == want
    assert want != (to_int(a) * to_int(b) >> (64 * lo)) % (1 << 320)
