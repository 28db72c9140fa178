"""Writes the golden fixtures under tests/golden/.

The reference (FLINT) cannot be built here from its own sources without its
cmake configure step (generated flint.h / flint-config.h) and GMP, so the
exact products are produced by GMP's mpn_mul -- the independent multiplier
FLINT's own t-mul.c:54 checks fft_small against -- and, for polynomial
products, by Kronecker substitution on Python integers.  Inputs come from
the seeded generator (SplitMix64 -> xoshiro256**) of oracle/fsmall_oracle.c.

    python tests/golden/make_fixtures.py

Outputs:
  mpn_small.npz     inputs and full products for a sweep of small shapes
  checksums.json    SHA-256 and mod-q residues of the config-size products
  poly_small.npz    nmod and fmpz products (inputs, moduli, outputs)
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

Q_CHECK = [2305843009213693951, 1152921504606846883, 4611686018427387847]


def gmp():
    for name in ("libgmp.so.10", "/usr/lib/x86_64-linux-gnu/libgmp.so.10"):
        try:
            L = ctypes.CDLL(name)
            f = L.__gmpn_mul
            f.restype = ctypes.c_uint64
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                          ctypes.c_long]
            return f
        except OSError:
            continue
    return None


def gmp_mul(a, b):
    f = gmp()
    if f is None:
        raise RuntimeError("libgmp not available")
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    if len(a) < len(b):
        a, b = b, a
    z = np.zeros(len(a) + len(b), dtype=np.uint64)
    f(z.ctypes.data, a.ctypes.data, len(a), b.ctypes.data, len(b))
    return z


def rng_pair(seed, an, bn):
    """both operands from one stream: a then b, top limbs nonzero"""
    r = O.lib()
    st = (ctypes.c_uint64 * 4)()
    r.or_rng_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    r.or_rng_limbs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
    r.or_rng_seed(st, seed)
    a = np.empty(an, dtype=np.uint64)
    b = np.empty(bn, dtype=np.uint64)
    r.or_rng_limbs(st, a.ctypes.data, an, 1)
    r.or_rng_limbs(st, b.ctypes.data, bn, 1)
    return a, b


def checksum(z):
    zb = np.ascontiguousarray(z, dtype=np.uint64).tobytes()
    x = int.from_bytes(zb, "little")
    return {"sha256": hashlib.sha256(zb).hexdigest(), "mod_q": [str(x % q) for q in Q_CHECK],
            "limbs": len(z)}


SMALL_SHAPES = [(1, 1), (2, 1), (5, 3), (17, 16), (100, 37), (256, 256), (511, 257),
                (1000, 999), (2048, 1), (3000, 2500), (5000, 4999), (8191, 4097)]


def poly_ks(a, b, signed=False):
    """exact coefficient product by Kronecker substitution on Python ints"""
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    ma = max((abs(x) for x in a), default=0)
    mb = max((abs(x) for x in b), default=0)
    bound = ma * mb * min(len(a), len(b)) + 1
    slot = bound.bit_length() + 2
    base = 1 << slot

    def pack(v):
        x = 0
        for c in reversed(v):
            x = x * base + c
        return x
    prod = pack(a) * pack(b)
    out = []
    for _ in range(len(a) + len(b) - 1):
        r = prod % base
        if signed and r >= base // 2:
            r -= base
        out.append(r)
        prod = (prod - r) // base
    return out


def main():
    rec = {}
    # small mpn products (full vectors)
    arrays = {}
    for i, (an, bn) in enumerate(SMALL_SHAPES):
        a, b = rng_pair(0xF1A70000 + 100 + i, an, bn)
        z = gmp_mul(a, b)
        assert int.from_bytes(z.tobytes(), "little") == \
            int.from_bytes(a.tobytes(), "little") * int.from_bytes(b.tobytes(), "little")
        arrays["a%d" % i] = a
        arrays["b%d" % i] = b
        arrays["z%d" % i] = z
    arrays["shapes"] = np.array(SMALL_SHAPES, dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "mpn_small.npz"), **arrays)

    # config-size products: checksums only
    for name, seed, n in (("config1_65536", 0xF1A70001, 65536),
                          ("config2_2p24", 0xF1A70002, 1 << 24)):
        a, b = rng_pair(seed, n, n)
        z = gmp_mul(a, b)
        rec[name] = dict(seed=seed, an=n, bn=n, **checksum(z))
        rec[name]["a_sha256"] = hashlib.sha256(a.tobytes()).hexdigest()
        print(name, rec[name]["sha256"])
    rec["q_check"] = [str(q) for q in Q_CHECK]
    with open(os.path.join(HERE, "checksums.json"), "w") as f:
        json.dump(rec, f, indent=1)

    # polynomial products
    rs = np.random.default_rng(0xF1A70003)
    P = {}
    cases = []
    moduli = [2, 3, 7, 257, 65537, 7340033, 1108307720798209, 659706976665601,
              (1 << 61) - 1, (1 << 63) + 29, (1 << 64) - 59, 1000003, (1 << 50) - 27]
    for k, n in enumerate(moduli):
        an = int(rs.integers(1, 3000))
        bn = int(rs.integers(1, 3000))
        a = [int(x) % n for x in rs.integers(0, 2**63, an, dtype=np.uint64)]
        b = [int(x) % n for x in rs.integers(0, 2**63, bn, dtype=np.uint64)]
        zfull = [c % n for c in poly_ks(a, b)]
        P["nmod_a%d" % k] = np.array(a, dtype=np.uint64)
        P["nmod_b%d" % k] = np.array(b, dtype=np.uint64)
        P["nmod_z%d" % k] = np.array(zfull, dtype=np.uint64)
        cases.append(n)
    P["nmod_moduli"] = np.array([str(n) for n in cases])
    fcases = []
    for k, (bits, signed) in enumerate([(10, False), (40, True), (64, True), (100, True),
                                        (150, False), (190, True)]):
        an = int(rs.integers(1, 1500))
        bn = int(rs.integers(1, 1500))
        w = (bits + 64) // 64
        zw = (2 * bits + 24 + 64) // 64 + 1

        def draw(n):
            out = []
            for _ in range(n):
                x = int.from_bytes(rs.bytes((bits + 7) // 8), "little") % (1 << bits)
                if signed and rs.integers(0, 2):
                    x = -x
                out.append(x)
            return out
        a = draw(an)
        b = draw(bn)
        zc = poly_ks(a, b, signed=True)

        def enc(v, width):
            return np.frombuffer(b"".join((x % (1 << (64 * width))).to_bytes(8 * width, "little")
                                          for x in v), dtype=np.uint64)
        P["fmpz_a%d" % k] = enc(a, w)
        P["fmpz_b%d" % k] = enc(b, w)
        P["fmpz_z%d" % k] = enc(zc, zw)
        fcases.append((w, w, zw))
    P["fmpz_widths"] = np.array(fcases, dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "poly_small.npz"), **P)
    print("fixtures written")


if __name__ == "__main__":
    main()
