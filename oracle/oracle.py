"""ctypes wrapper for the CPU oracle (oracle/libfsmall_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker of the GPU product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfsmall_oracle.so")
_lib = None

u64 = ctypes.c_uint64
s64 = ctypes.c_int64
P64 = ctypes.POINTER(ctypes.c_uint64)


class NmodT(ctypes.Structure):
    _fields_ = [("n", u64), ("ninv", u64), ("norm", u64)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_is_prime.argtypes = [u64]
        L.or_is_prime.restype = ctypes.c_int
        L.or_next_fft_number.argtypes = [u64]
        L.or_next_fft_number.restype = u64
        L.or_default_primes.argtypes = [P64]
        L.or_mulmod_satisfies_bounds.argtypes = [u64]
        L.or_mulmod_satisfies_bounds.restype = ctypes.c_int
        for f in ("or_fp_mulmod",):
            getattr(L, f).argtypes = [ctypes.c_double] * 4
            getattr(L, f).restype = ctypes.c_double
        for f in ("or_fp_reduce_to_pm1n", "or_fp_reduce_to_0n"):
            getattr(L, f).argtypes = [ctypes.c_double] * 3
            getattr(L, f).restype = ctypes.c_double
        L.or_quadratic_nonresidue.argtypes = [u64]
        L.or_quadratic_nonresidue.restype = u64
        L.or_primitive_2power_root.argtypes = [u64]
        L.or_primitive_2power_root.restype = u64
        L.or_w.argtypes = [u64, u64]
        L.or_w.restype = u64
        L.or_w2.argtypes = [u64, u64]
        L.or_w2.restype = u64
        L.or_trunc_index.argtypes = [u64, u64]
        L.or_trunc_index.restype = u64
        L.or_fft_trunc.argtypes = [u64, P64, u64, u64, u64]
        L.or_fft_trunc.restype = ctypes.c_int
        L.or_ifft_trunc.argtypes = [u64, P64, u64, u64]
        L.or_ifft_trunc.restype = ctypes.c_int
        L.or_crt_bn_bound.argtypes = [ctypes.c_int, u64]
        L.or_crt_bn_bound.restype = u64
        L.or_crt_find_bits.argtypes = [ctypes.c_int, u64]
        L.or_crt_find_bits.restype = u64
        L.or_best_profile.argtypes = [u64, u64, u64, P64, P64]
        L.or_best_profile.restype = ctypes.c_int
        L.or_np2_choose.argtypes = [u64, u64, u64, P64, P64]
        L.or_np2_choose.restype = ctypes.c_int
        L.or_mpn_plan.argtypes = [u64, u64, u64, P64]
        L.or_mpn_plan.restype = ctypes.c_int
        L.or_mpn_to_residues.argtypes = [u64, P64, u64, P64, u64, u64]
        L.or_mpn_mul.argtypes = [P64, P64, u64, P64, u64, u64]
        L.or_mpn_mul.restype = ctypes.c_int
        L.or_mpn_mul_basecase.argtypes = [P64, P64, u64, P64, u64]
        L.or_mpn_mul_range.argtypes = [P64, u64, u64, P64, u64, P64, u64, u64]
        L.or_mpn_mul_range.restype = ctypes.c_int
        L.or_nmod_poly_mul_mid.argtypes = [P64, s64, s64, P64, s64, P64, s64, NmodT]
        L.or_nmod_poly_mul_mid.restype = ctypes.c_int
        L.or_fmpz_poly_mul_mid.argtypes = [P64, u64, s64, s64, P64, u64, s64, P64, u64, s64]
        L.or_fmpz_poly_mul_mid.restype = ctypes.c_int
        L.or_fmpz_poly_nprimes.argtypes = [P64, u64, s64, P64, u64, s64]
        L.or_fmpz_poly_nprimes.restype = ctypes.c_int
        L.or_rng_seed.argtypes = [ctypes.c_void_p, u64]
        L.or_rng_next.argtypes = [ctypes.c_void_p]
        L.or_rng_next.restype = u64
        L.or_rng_uniform.argtypes = [ctypes.c_void_p, u64, u64]
        L.or_rng_uniform.restype = u64
        L.or_rng_random_prime.argtypes = [ctypes.c_void_p, u64]
        L.or_rng_random_prime.restype = u64
        L.or_rng_limbs.argtypes = [ctypes.c_void_p, P64, u64, ctypes.c_int]
        L.or_rng_double.argtypes = [ctypes.c_void_p]
        L.or_rng_double.restype = ctypes.c_double
        L.or_rng_log_uniform.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.or_rng_log_uniform.restype = u64
        L.or_rng_fill_below.argtypes = [ctypes.c_void_p, P64, u64, u64]
        L.or_rng_fill_signed.argtypes = [ctypes.c_void_p, P64, u64, ctypes.c_int, u64]
        L.or_mpn_mod_q.argtypes = [P64, u64, u64]
        L.or_mpn_mod_q.restype = u64
        L.or_nmod_poly_eval.argtypes = [P64, u64, u64, u64]
        L.or_nmod_poly_eval.restype = u64
        L.or_fmpz_poly_eval_mod_q.argtypes = [P64, u64, u64, u64, u64]
        L.or_fmpz_poly_eval_mod_q.restype = u64
        L.or_stream_word.argtypes = [u64, u64]
        L.or_stream_word.restype = u64
        L.or_stream_headers.argtypes = [u64, u64, P64]
        L.or_stream_fill.argtypes = [ctypes.c_int, u64, u64, u64, u64, ctypes.c_int, P64]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(P64)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def default_primes():
    out = np.zeros(8, dtype=np.uint64)
    lib().or_default_primes(_ptr(out))
    return [int(x) for x in out]


def fft_trunc(p, data, L, itrunc, otrunc):
    d = _u64(data).copy()
    assert len(d) >= (1 << L)
    assert lib().or_fft_trunc(p, _ptr(d), L, itrunc, otrunc)
    return d


def ifft_trunc(p, data, L, trunc):
    d = _u64(data).copy()
    assert lib().or_ifft_trunc(p, _ptr(d), L, trunc)
    return d


def mpn_plan(an, bn, nthreads=1):
    out = np.zeros(6, dtype=np.uint64)
    assert lib().or_mpn_plan(an, bn, nthreads, _ptr(out))
    return dict(zip(("np", "bits", "depth", "ztrunc", "alen", "blen"), (int(x) for x in out)))


def best_profile(an, bn, nthreads=1):
    np_ = u64()
    bits = u64()
    lib().or_best_profile(an, bn, nthreads, ctypes.byref(np_), ctypes.byref(bits))
    return int(np_.value), int(bits.value)


def mpn_to_residues(p, a, bits, dn):
    a = _u64(a)
    d = np.zeros(dn, dtype=np.uint64)
    lib().or_mpn_to_residues(p, _ptr(d), dn, _ptr(a), len(a), bits)
    return d


def mpn_mul(a, b, nthreads=1):
    a = _u64(a)
    b = _u64(b)
    z = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert lib().or_mpn_mul(_ptr(z), _ptr(a), len(a), _ptr(b), len(b), nthreads)
    return z


def mpn_mul_range(a, b, lo, hi, nthreads=1):
    """limbs [lo, hi) as _mpn_ctx_mpn_mul_range returns them (FLINT on
    nthreads threads)"""
    a = _u64(a)
    b = _u64(b)
    z = np.zeros(max(hi - lo, 1), dtype=np.uint64)
    assert lib().or_mpn_mul_range(_ptr(z), lo, hi, _ptr(a), len(a), _ptr(b), len(b), nthreads)
    return z[:hi - lo]


def mpn_mul_basecase(a, b):
    a = _u64(a)
    b = _u64(b)
    z = np.zeros(len(a) + len(b), dtype=np.uint64)
    lib().or_mpn_mul_basecase(_ptr(z), _ptr(a), len(a), _ptr(b), len(b))
    return z


def nmod_init(n):
    norm = 64 - int(n).bit_length()
    ninv = ((1 << 128) - 1) // (n << norm) - (1 << 64) if n else 0
    return NmodT(n, ninv & ((1 << 64) - 1), norm)


def nmod_poly_mul_mid(a, b, zl, zh, n):
    a = _u64(a)
    b = _u64(b)
    z = np.zeros(max(zh - zl, 1), dtype=np.uint64)
    assert lib().or_nmod_poly_mul_mid(_ptr(z), zl, zh, _ptr(a), len(a), _ptr(b), len(b),
                                      nmod_init(n))
    return z[: zh - zl]


def fmpz_poly_mul_mid(a, aw, b, bw, zl, zh, zw):
    """a, b: flat uint64 arrays of an*aw, bn*bw limbs (two's complement).
    Returns (ok, z) with z of (zh-zl)*zw limbs, or (0, None) when declined."""
    a = _u64(a)
    b = _u64(b)
    an = len(a) // aw
    bn = len(b) // bw
    z = np.zeros(max((zh - zl) * zw, 1), dtype=np.uint64)
    ok = lib().or_fmpz_poly_mul_mid(_ptr(z), zw, zl, zh, _ptr(a), aw, an, _ptr(b), bw, bn)
    return ok, (z[: (zh - zl) * zw] if ok else None)


def fmpz_poly_nprimes(a, aw, b, bw):
    a = _u64(a)
    b = _u64(b)
    return lib().or_fmpz_poly_nprimes(_ptr(a), aw, len(a) // aw, _ptr(b), bw, len(b) // bw)


# ----------------------------------------------------------------------
# full-size checkers: residues mod a word-size q
# ----------------------------------------------------------------------

def mpn_mod_q(a, q):
    """the integer with little-endian limbs a, mod q"""
    a = _u64(a)
    return int(lib().or_mpn_mod_q(_ptr(a), len(a), q))


def nmod_poly_eval(c, x, q):
    """sum c_i x^i mod q"""
    c = _u64(c)
    return int(lib().or_nmod_poly_eval(_ptr(c), len(c), x, q))


def fmpz_poly_eval_mod_q(c, w, x, q):
    """sum c_i x^i mod q over w-limb two's complement coefficients"""
    c = _u64(c)
    return int(lib().or_fmpz_poly_eval_mod_q(_ptr(c), w, len(c) // w, x, q))


# ----------------------------------------------------------------------
# the batched workloads' request generator, restated (SplitMix64 seeding a
# xoshiro256** stream, seed 0xF1A70000 + config index; draws in this order)
# ----------------------------------------------------------------------

class Rng:
    def __init__(self, seed):
        self.s = (ctypes.c_uint64 * 4)()
        lib().or_rng_seed(self.s, seed)

    def uniform(self, lo, hi):
        return int(lib().or_rng_uniform(self.s, lo, hi))

    def double(self):
        return float(lib().or_rng_double(self.s))

    def log_uniform(self, x, y):
        return int(lib().or_rng_log_uniform(self.s, x, y))

    def prime(self, bits):
        return int(lib().or_rng_random_prime(self.s, bits))

    def below(self, n, bound):
        out = np.zeros(n, dtype=np.uint64)
        lib().or_rng_fill_below(self.s, _ptr(out), n, bound)
        return out

    def limbs(self, n):
        out = np.zeros(n, dtype=np.uint64)
        lib().or_rng_limbs(self.s, _ptr(out), n, 1)
        return out

    def signed(self, n, bits, w):
        out = np.zeros(n * w, dtype=np.uint64)
        lib().or_rng_fill_signed(self.s, _ptr(out), n, bits, w)
        return out


def workload_requests(workload, count):
    """the first `count` requests of a workload: tuples
    ("nmod", a, b, n) / ("fmpz", a, b, w, bits) / ("mpn", a, b)"""
    seed = {"nmod4096": 0xF1A70003, "fmpz256": 0xF1A70004, "stream": 0xF1A70005}[workload]
    r = Rng(seed)

    def nmod():
        an, bn = r.log_uniform(6, 16), r.log_uniform(6, 16)
        n = r.prime(r.uniform(20, 63))
        return ("nmod", r.below(an, n), r.below(bn, n), n)

    def fmpz():
        an, bn = r.uniform(1024, 16384), r.uniform(1024, 16384)
        bits = r.uniform(64, 4096)
        w = bits // 64 + 1
        return ("fmpz", r.signed(an, bits, w), r.signed(bn, bits, w), w, bits)

    def mpn():
        an, bn = r.log_uniform(4, 22), r.log_uniform(4, 22)
        return ("mpn", r.limbs(an), r.limbs(bn))

    out = []
    for _ in range(count):
        if workload == "nmod4096":
            out.append(nmod())
        elif workload == "fmpz256":
            out.append(fmpz())
        else:
            u = r.double()
            out.append(mpn() if u < 0.5 else nmod() if u < 0.8 else fmpz())
    return out


# ----------------------------------------------------------------------
# the full mixed stream: headers, and operand contents keyed per operand
# ----------------------------------------------------------------------

def stream_word(key, j):
    return int(lib().or_stream_word(key, j))


def stream_headers(count, seed=0xF1A70005):
    """(count, 7) uint64: kind (0 integer, 1 nmod, 2 fmpz), bits, an, bn, n,
    key0, key1"""
    out = np.zeros((count, 7), dtype=np.uint64)
    lib().or_stream_headers(seed, count, _ptr(out))
    return out


def stream_operand(kind, key, count, n=0, w=1, bits=0):
    out = np.zeros(count * (w if kind == 2 else 1), dtype=np.uint64)
    lib().or_stream_fill(kind, key, count, n, w, bits, _ptr(out))
    return out
