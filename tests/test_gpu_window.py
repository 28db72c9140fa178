"""GPU parity of windowed and middle products.

- Integer windows (_mpn_ctx_mpn_mul_range, flint_mpn_mulmid_fft_small):
  the reference's words, a lower approximation whose dropped carry follows
  FLINT's plan (per chunk on two-prime plans, below coefficient c_lo on
  wide plans), against the oracle's restatement (or_mpn_mul_range) on
  random shapes and windows, a planted wide-plan dropped carry, other FLINT
  thread counts and the exact window (nthreads = 0); every window also
  meets t-mulmid.c's deficit bound.  Above the oracle's sizes the words are
  checked against the exact window and the correction computed on the host
  from the boundary coefficients (tests/test_window_plan.py's model, itself
  checked against the oracle there).
- Polynomial windows with wraparound (plan.c:47-89): when the operands and
  zh fit in 2^floor(log2 zn) and zn - 2^d <= zl, the product is taken mod
  X^(2^d) - 1 at half the transform length.  Middle products of that shape
  (nmod and fmpz) equal the oracle's window, and a batch request reports
  the shorter transform.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import from_int, rand_limbs, to_int
from tests.test_oracle_window import _coeff_len
from tests.test_window_plan import _model_window

pytestmark = pytest.mark.gpu


def _deficit_ok(full, z, zlo, zhi, an, bn):
    zn = zhi - zlo
    d = (to_int(full[zlo:zhi]) - to_int(z)) % (1 << (64 * zn))
    if zlo == 0:
        return d == 0
    if zn >= 2:
        return (d >> 128) == 0 and ((d >> 64) & ((1 << 64) - 1)) <= an + bn
    return True


def test_mpn_windows_match_reference_words():
    import flint_amd as fa
    rng = np.random.default_rng(2024)
    shapes = [(int(rng.integers(1, 31)), int(rng.integers(1, 31))) for _ in range(60)]
    shapes += [(600, 17), (5000, 4000), (40000, 39999), (40001, 39999), (70000, 65536),
               (45000, 3)]
    lowered = 0
    for an, bn in shapes:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        full = O.mpn_mul(a, b)
        for _ in range(4):
            zlo = int(rng.integers(0, an + bn))
            zhi = zlo + 1 + int(rng.integers(0, an + bn - zlo))
            print(an, bn, zlo, zhi, flush=True)
            z = fa.mpn_mul_range(a, b, zlo, zhi)
            assert np.array_equal(z, O.mpn_mul_range(a, b, zlo, zhi)), (an, bn, zlo, zhi)
            assert _deficit_ok(full, z, zlo, zhi, an, bn), (an, bn, zlo, zhi)
            lowered += not np.array_equal(z, full[zlo:zhi])
    assert lowered > 5      # two-prime windows drop chunk tails' carries often


def test_mpn_window_edges():
    """window edges against the oracle: one limb at either end, everything
    above limb 1, and windows starting just below, at and above the limb
    count n of the wide plans' coefficients (c_lo = 0 up to lo = n)"""
    import flint_amd as fa
    rng = np.random.default_rng(99)
    for an, bn in ((7, 3), (3, 7), (1000, 999), (39999, 40000), (50000, 7), (7, 50000),
                   (60000, 41000)):
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        zn = an + bn
        wins = [(0, 1), (zn - 1, zn), (1, zn), (zn // 2, zn // 2 + 1)]
        wins += [(lo, min(lo + 3, zn)) for lo in range(2, 10) if lo < zn]
        for lo, hi in wins:
            z = fa.mpn_mul_range(a, b, lo, hi)
            assert np.array_equal(z, O.mpn_mul_range(a, b, lo, hi)), (an, bn, lo, hi)


def test_mpn_window_planted_dropped_carry():
    """wide plan (an > 40000), b = 2^64 - 1: a's bits in [c_lo bits, 64 lo)
    make the coefficients from c_lo up sum to T - 2^(c_lo bits) below T, so
    the dropped coefficients carry into limb lo: one below the exact window"""
    import flint_amd as fa
    rng = np.random.default_rng(3)
    for an, lo in ((50000, 30011), (41000, 777), (300000, 123457)):
        a = rand_limbs(rng, an)
        b = np.array([2**64 - 1], dtype=np.uint64)
        pl = O.mpn_plan(an, 1)
        bits, nn = pl["bits"], _coeff_len(pl["np"])
        c_lo = -(-(lo - nn) * 64 // bits)
        r = 64 * lo - c_lo * bits
        m = (-pow(2**64 - 1, -1, 1 << r)) % (1 << r)
        x = (to_int(a) & ~(((1 << r) - 1) << (c_lo * bits))) | (m << (c_lo * bits))
        a = from_int(x, an)
        exact = (x * (2**64 - 1)) >> (64 * lo)
        for hi in (lo + 1, lo + 40, an + 1):
            print(an, lo, hi, flush=True)
            z = fa.mpn_mul_range(a, b, lo, hi)
            w = 1 << (64 * (hi - lo))
            assert to_int(z) == (exact - 1) % w, (an, lo, hi)
            if an < 100000:
                assert np.array_equal(z, O.mpn_mul_range(a, b, lo, hi))


def test_mpn_windows_large_against_model():
    """sizes above the oracle's: the exact product's window minus the
    correction computed on the host from the boundary coefficients"""
    import flint_amd as fa
    rng = np.random.default_rng(77)
    cases = {(1 << 20, 3 << 18): [(17, 4000), (300, 1 << 19)],
             (300000, 299999): [(299990, 300100)],
             (1 << 21, 5000): [(int(x), int(x) + 64) for x in rng.integers(1, 1 << 21, 3)]}
    for (an, bn), wins in cases.items():
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        full = fa.mpn_mul(a, b)       # exact (pinned against GMP elsewhere)
        for lo, hi in wins:
            print(an, bn, lo, hi, flush=True)
            z = fa.mpn_mul_range(a, b, lo, hi)
            assert to_int(z) == _model_window(a, b, lo, hi, full=full), (an, bn, lo, hi)


def test_mpn_window_thread_counts_and_exact():
    """device operands: FLINT on 2 and 8 threads (other plans, other
    words), and nthreads = 0 for the exact window"""
    import torch
    import flint_amd as fa
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(21)
    for an, bn in ((60000, 50000), (1500, 700), (45000, 300)):
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        full = O.mpn_mul(a, b)
        da = torch.from_numpy(a.view(np.int64)).to(dev)
        db = torch.from_numpy(b.view(np.int64)).to(dev)
        torch.cuda.synchronize()
        s = torch.cuda.Stream(device=dev)
        for lo, hi in ((an // 3, an // 3 + 50), (an - 2, an + bn)):
            for nt in (0, 1, 2, 3, 5, 6, 7, 8):
                print(an, bn, lo, hi, nt, flush=True)
                dz = torch.zeros(hi - lo, dtype=torch.int64, device=dev)
                torch.cuda.synchronize()
                fa.mpn_mul_range_threads(dz.data_ptr(), lo, hi, da.data_ptr(), an, db.data_ptr(),
                                         bn, nt, s.cuda_stream)
                s.synchronize()
                got = dz.cpu().numpy().view(np.uint64)
                want = full[lo:hi] if nt == 0 else O.mpn_mul_range(a, b, lo, hi, nt)
                assert np.array_equal(got, want), (an, bn, lo, hi, nt)
                # the host-pointer entry with the same thread count
                assert np.array_equal(fa.mpn_mul_range(a, b, lo, hi, nt), want)


def test_mpn_middle_product_device():
    """the balanced middle product of {a, 2n-1} and {b, n}: limbs
    [n-1, 2n+1) (flint_mpn_mulmid_n's band) on device pointers, the
    reference's words"""
    import torch
    import flint_amd as fa
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(8)
    for n in (100, 9000, 30000):
        a, b = rand_limbs(rng, 2 * n - 1), rand_limbs(rng, n)
        want = O.mpn_mul_range(a, b, n - 1, 2 * n + 1)
        da = torch.from_numpy(a.view(np.int64)).to(dev)
        db = torch.from_numpy(b.view(np.int64)).to(dev)
        dz = torch.zeros(n + 2, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()   # the fill runs on torch's stream, the call on s
        s = torch.cuda.Stream(device=dev)
        fa.mpn_mul_range_device(dz.data_ptr(), n - 1, 2 * n + 1, da.data_ptr(), 2 * n - 1,
                                db.data_ptr(), n, s.cuda_stream)
        s.synchronize()
        assert np.array_equal(dz.cpu().numpy().view(np.uint64), want), n


@pytest.mark.parametrize("n", [(1 << 61) - 1, 65537, 2**64 - 59])
@pytest.mark.parametrize("m", [50, 3000, 70000])
def test_nmod_middle_products_wrap(n, m):
    """a of 2m - 1 coefficients, b of m: window [m - 1, 2m - 1) of a
    product of 3m - 2 coefficients, computed mod X^(2^d) - 1"""
    import flint_amd as fa
    rng = np.random.default_rng(m)
    if n < 2**63:
        a = rng.integers(0, n, 2 * m - 1, dtype=np.uint64)
        b = rng.integers(0, n, m, dtype=np.uint64)
    else:
        a = rng.integers(0, 2**63, 2 * m - 1, dtype=np.uint64)
        b = rng.integers(0, 2**63, m, dtype=np.uint64)
    zn = 3 * m - 2
    zl, zh = m - 1, 2 * m - 1
    want = O.nmod_poly_mul_mid(a, b, zl, zh, n)
    assert np.array_equal(fa.nmod_poly_mul_mid(a, b, zl, zh, n), want)
    # a window that does not allow the wrap is computed at full length
    assert np.array_equal(fa.nmod_poly_mul_mid(a, b, 0, zh, n),
                          O.nmod_poly_mul_mid(a, b, 0, zh, n))
    # the batch reports the depth it used: 2^floor(log2 zn) when the wrap applies
    d = zn.bit_length() - 1
    wraps = max(len(a), len(b)) <= (1 << d) and zh <= (1 << d) <= zn <= zl + (1 << d)
    R = (fa.Request * 1)()
    z = np.zeros(zh - zl, dtype=np.uint64)
    R[0].kind = fa.REQ_NMOD
    R[0].a, R[0].b, R[0].z = a.ctypes.data, b.ctypes.data, z.ctypes.data
    R[0].an, R[0].bn, R[0].zl, R[0].zh = len(a), len(b), zl, zh
    R[0].mod = fa.nmod_init(n)
    assert fa.lib().fa_batch_mul(ctypes.cast(R, ctypes.c_void_p), 1, 0) == 0
    assert R[0].status == 1 and np.array_equal(z, want)
    want_depth = max(4, d if wraps else (zn - 1).bit_length())
    assert R[0].depth == want_depth, (R[0].depth, want_depth, wraps)


def test_fmpz_middle_products_wrap():
    import flint_amd as fa
    rng = np.random.default_rng(31)
    w, zw = 2, 4
    for m in (40, 2500):
        av = [int(v) for v in rng.integers(-2**62, 2**62, 2 * m - 1)]
        bv = [int(v) for v in rng.integers(-2**62, 2**62, m)]

        def pack(vals):
            out = np.zeros(len(vals) * w, dtype=np.uint64)
            for i, v in enumerate(vals):
                out[i * w:(i + 1) * w] = np.frombuffer((v % (1 << 128)).to_bytes(16, "little"),
                                                       dtype=np.uint64)
            return out
        a, b = pack(av), pack(bv)
        zl, zh = m - 1, 2 * m - 1
        got = fa.fmpz_poly_mul_mid(a, w, b, w, zl, zh, zw)
        ok, want = O.fmpz_poly_mul_mid(a, w, b, w, zl, zh, zw)
        assert ok and got is not None and np.array_equal(got, want), m
