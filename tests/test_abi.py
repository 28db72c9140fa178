"""CPU tests of the C ABI: the library loads, exports every entry point the
header declares, and its host-only helpers behave (no GPU compute here)."""
import ctypes
import os
import re

import numpy as np

import flint_amd as fa
from oracle import oracle as O
from tests.golden.make_fixtures import rng_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "flint_amd.h")).read()
    return re.findall(r"^FA_API [^(]*?\b(\w+)\(", src, flags=re.M)


def test_library_exports_every_declared_symbol():
    L = fa.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in fa.SIGNATURES, name


def test_drop_in_symbols_have_reference_names():
    L = fa.lib()
    for name in ("mpn_mul_default_mpn_ctx", "_nmod_poly_mul_mid_default_mpn_ctx"):
        assert hasattr(L, name)


def test_default_primes_host_side():
    assert fa.default_primes() == O.default_primes()


def test_synthetic_generator_matches_oracle_generator():
    for seed, an, bn in ((0xF1A70002, 1000, 777), (1, 5, 1), (0xF1A70001, 4096, 4096)):
        a, b = fa.synth_mpn_pair(seed, an, bn)
        ra, rb = rng_pair(seed, an, bn)
        assert np.array_equal(a, ra) and np.array_equal(b, rb)
        assert a[-1] != 0 and b[-1] != 0


def test_gpu_plan_host_side():
    p = fa.mpn_plan(1 << 24, 1 << 24)
    assert (p["np"], p["bits"], p["depth"]) == (6, 136, 24)
    assert p["zlen"] <= 1 << 24
    p1 = fa.mpn_plan(65536, 65536)
    assert p1["depth"] in (16, 17) and p1["bits"] % 8 == 0
    # exactness bound of every plan: min(alen, blen) * 2^(2 bits) <= prod p_i
    primes = O.default_primes()
    for an, bn in ((1, 1), (7, 3), (1000, 10), (40000, 40000), (3 << 20, 5 << 20)):
        p = fa.mpn_plan(an, bn)
        M = 1
        for q in primes[:p["np"]]:
            M *= q
        alen = -(-64 * an // p["bits"])
        blen = -(-64 * bn // p["bits"])
        assert min(alen, blen) << (2 * p["bits"]) <= M
        assert alen + blen - 1 <= 1 << p["depth"]


def test_errors_are_raised_not_swallowed():
    """argument errors come back as exceptions from the Python mirror"""
    L = fa.lib()
    rc = L.fa_mpn_plan(0, 5, (ctypes.c_uint64 * 6)())
    assert rc < 0
    assert b"an, bn" in L.fa_last_error()
    try:
        fa.mpn_plan(-1, 3)
        assert False, "no exception"
    except fa.FlintAmdError:
        pass


def test_mulmod_bounds_check_matches_oracle():
    """fft_small_mulmod_satisfies_bounds (host side) against the oracle's
    restatement, on the default primes and moduli around 2^50"""
    rng = np.random.default_rng(9)
    cands = [int(p) for p in O.default_primes()] + [0x0003f00000000001, 257, 65537]
    cands += [int(x) | 1 for x in rng.integers(1 << 40, 1 << 52, 200)]
    for n in cands:
        assert fa.lib().fa_mulmod_satisfies_bounds(n) == O.lib().or_mulmod_satisfies_bounds(n), n
