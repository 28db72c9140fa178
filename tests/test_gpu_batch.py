"""GPU parity of the batched form: a batch of mixed integer, nmod and fmpz
products gives, request by request, the output of the corresponding
single-product entry point (itself pinned to the oracle and exact integers
in test_gpu_mpn.py / test_gpu_poly.py), and exact products where Python
integers can check them.  The device form is captured into a hipGraph and
replayed several times with eager work in between; every replay is checked.
"""
import numpy as np
import pytest

from tests.helpers import to_int, rand_limbs

pytestmark = pytest.mark.gpu


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def random_prime(rng, bits):
    while True:
        n = int(rng.integers(1 << (bits - 1), 1 << bits, dtype=np.uint64)) | 1
        if bits == 64:
            n = int.from_bytes(rng.bytes(8), "little") | (1 << 63) | 1
        if _is_prime(n):
            return n


def signed_coeffs(rng, n, bits, w):
    out = np.zeros(n * w, dtype=np.uint64)
    for i in range(n):
        x = int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits)
        if rng.integers(0, 2):
            x = -x
        v = x % (1 << (64 * w))
        out[i * w:(i + 1) * w] = np.frombuffer(v.to_bytes(8 * w, "little"), dtype=np.uint64)
    return out


def mixed_requests(rng):
    import flint_amd as fa
    reqs = []
    for an, bn in [(1, 1), (7, 3), (100, 37), (2000, 1900), (30000, 1000), (65536, 65536),
                   (100000, 90000), (5, 40000)]:
        reqs.append({"kind": "mpn", "a": rand_limbs(rng, an), "b": rand_limbs(rng, bn)})
    primes = fa.default_primes()
    moduli = [1, 2, 3, 257, 7340033, primes[3], (1 << 61) - 1, (1 << 64) - 59]
    moduli += [random_prime(rng, int(b)) for b in rng.integers(20, 64, 6)]
    for k, n in enumerate(moduli):
        an = int(rng.integers(1, 30000)) if k % 3 else int(rng.integers(1, 300))
        bn = int(rng.integers(1, 30000)) if k % 2 else int(rng.integers(1, 300))
        a = np.array([int(x) % n for x in rng.integers(0, 2**63, an, dtype=np.uint64)],
                     dtype=np.uint64)
        b = np.array([int(x) % n for x in rng.integers(0, 2**63, bn, dtype=np.uint64)],
                     dtype=np.uint64)
        zn = an + bn - 1
        zl, zh = [(0, zn), (zn // 3, zn // 3 + 17), (max(zn - 5, 0), zn + 9)][k % 3]
        reqs.append({"kind": "nmod", "a": a, "b": b, "n": n, "zl": zl, "zh": zh})
    for bits, an, bn in [(10, 50, 40), (63, 3000, 2000), (100, 1000, 1500), (150, 17, 5000),
                         (190, 300, 300), (400, 100, 100)]:
        w = bits // 64 + 1
        zw = (2 * bits + 16 + 64) // 64 + 1
        a = signed_coeffs(rng, an, bits, w)
        b = signed_coeffs(rng, bn, bits, w)
        zn = an + bn - 1
        reqs.append({"kind": "fmpz", "a": a, "aw": w, "b": b, "bw": w, "zl": 0, "zh": zn + 3,
                     "zw": zw})
    return reqs


def expected(q):
    import flint_amd as fa
    if q["kind"] == "mpn":
        return 1, fa.mpn_mul(q["a"], q["b"])
    if q["kind"] == "nmod":
        return 1, fa.nmod_poly_mul_mid(q["a"], q["b"], q["zl"], q["zh"], q["n"])
    z = fa.fmpz_poly_mul_mid(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"], q["zw"])
    return (0, None) if z is None else (1, z)


def test_batch_mixed_matches_single_calls():
    import flint_amd as fa
    rng = np.random.default_rng(0xBA7C)
    reqs = mixed_requests(rng)
    print("batch of", len(reqs), "requests", flush=True)
    got = fa.batch_mul(reqs)
    for k, q in enumerate(reqs):
        st, z = got[k]
        est, ez = expected(q)
        assert st == est, (k, q["kind"], st, est)
        if st == 1:
            assert np.array_equal(z, ez), (k, q["kind"])
    # exact integers for the integer products
    for k, q in enumerate(reqs):
        if q["kind"] == "mpn":
            assert to_int(got[k][1]) == to_int(q["a"]) * to_int(q["b"])


def test_batch_nmod_shaped_like_the_batched_workload():
    """a draw shaped like the batched nmod workload (lengths log-uniform in
    [2^6, 2^12], prime moduli of 20..63 bits) against the oracle"""
    import flint_amd as fa
    from oracle import oracle as O
    rng = np.random.default_rng(0xF1A70003)
    reqs = []
    for _ in range(48):
        an = int(2 ** rng.uniform(6, 12))
        bn = int(2 ** rng.uniform(6, 12))
        n = random_prime(rng, int(rng.integers(20, 64)))
        a = np.array([int(x) % n for x in rng.integers(0, 2**63, an, dtype=np.uint64)],
                     dtype=np.uint64)
        b = np.array([int(x) % n for x in rng.integers(0, 2**63, bn, dtype=np.uint64)],
                     dtype=np.uint64)
        reqs.append({"kind": "nmod", "a": a, "b": b, "n": n, "zl": 0, "zh": an + bn - 1})
    got = fa.batch_mul(reqs)
    for k, q in enumerate(reqs):
        st, z = got[k]
        assert st == 1
        want = O.nmod_poly_mul_mid(q["a"], q["b"], 0, q["zh"], q["n"])
        assert np.array_equal(z, want), k


def test_batch_device_graph_replays():
    """device buffers, hipGraph capture, three replays with eager work and
    synchronizations between them; every replay's outputs are checked"""
    import ctypes
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(0xD3B1)
    dev = torch.device("cuda:0")
    shapes = [(3, 2), (500, 400), (20000, 3000), (70000, 65000)]
    nmods = [(4000, 3000, random_prime(rng, 40)), (100, 9000, random_prime(rng, 62))]
    R = (fa.Request * (len(shapes) + len(nmods)))()
    keep, want = [], []
    for k, (an, bn) in enumerate(shapes):
        a = rand_limbs(rng, an)
        b = rand_limbs(rng, bn)
        ta = torch.from_numpy(a.view(np.int64)).to(dev)
        tb = torch.from_numpy(b.view(np.int64)).to(dev)
        tz = torch.zeros(an + bn, dtype=torch.int64, device=dev)
        keep += [ta, tb, tz]
        r = R[k]
        r.kind = fa.REQ_MPN
        r.a, r.b, r.z = ta.data_ptr(), tb.data_ptr(), tz.data_ptr()
        r.an, r.bn = an, bn
        want.append((tz, fa.mpn_mul(a, b)))
    for k, (an, bn, n) in enumerate(nmods):
        a = np.array([int(x) % n for x in rng.integers(0, 2**63, an, dtype=np.uint64)],
                     dtype=np.uint64)
        b = np.array([int(x) % n for x in rng.integers(0, 2**63, bn, dtype=np.uint64)],
                     dtype=np.uint64)
        ta = torch.from_numpy(a.view(np.int64)).to(dev)
        tb = torch.from_numpy(b.view(np.int64)).to(dev)
        tz = torch.zeros(an + bn - 1, dtype=torch.int64, device=dev)
        keep += [ta, tb, tz]
        r = R[len(shapes) + k]
        r.kind = fa.REQ_NMOD
        r.a, r.b, r.z = ta.data_ptr(), tb.data_ptr(), tz.data_ptr()
        r.an, r.bn = an, bn
        r.zl, r.zh = 0, an + bn - 1
        r.mod = fa.nmod_init(n)
        want.append((tz, fa.nmod_poly_mul_mid(a, b, 0, an + bn - 1, n)))
    stream = torch.cuda.Stream()
    B = fa.DeviceBatch(R, graph=True)
    for st in [R[k].status for k in range(len(R))]:
        assert st == 1
    try:
        for rep in range(3):
            for tz, _ in want:
                tz.zero_()
            x = torch.randn(1 << 20, device=dev)
            (x * 2).sum().item()        # unrelated eager work between replays
            torch.cuda.synchronize()
            B.launch(ctypes.c_void_p(stream.cuda_stream))
            stream.synchronize()
            for k, (tz, w) in enumerate(want):
                got = tz.cpu().numpy().view(np.uint64)
                assert np.array_equal(got, w), (rep, k)
    finally:
        B.close()


def test_batch_fmpz_kronecker_route():
    """fmpz products beyond eight primes inside a batch (FA_BATCH_FMPZ_KS):
    the same outputs as the single-call Kronecker route, and exact on a
    small case; host and graph-captured device forms"""
    import ctypes
    import torch
    import flint_amd as fa
    from tests.golden.make_fixtures import poly_ks
    rng = np.random.default_rng(0x4B5)
    reqs = []
    for bits, an, bn in [(300, 40, 30), (700, 500, 400), (150, 800, 900), (2000, 64, 100)]:
        w = bits // 64 + 1
        zw = (2 * bits + 16 + 64) // 64 + 1
        a = signed_coeffs(rng, an, bits, w)
        b = signed_coeffs(rng, bn, bits, w)
        reqs.append({"kind": "fmpz", "a": a, "aw": w, "b": b, "bw": w, "zl": 0,
                     "zh": an + bn + 1, "zw": zw})
    got = fa.batch_mul(reqs, ks=True)
    for k, q in enumerate(reqs):
        st, z = got[k]
        assert st == 1, k
        want = fa.fmpz_poly_mul_ks(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"], q["zw"])
        assert np.array_equal(z, want), k
    # exact integers for the first one
    q = reqs[0]

    def dec(v, w):
        out = []
        for i in range(len(v) // w):
            x = int.from_bytes(np.ascontiguousarray(v[i * w:(i + 1) * w]).tobytes(), "little")
            out.append(x - (1 << (64 * w)) if x >> (64 * w - 1) else x)
        return out
    exact = poly_ks(dec(q["a"], q["aw"]), dec(q["b"], q["bw"]), signed=True)
    assert dec(got[0][1], q["zw"])[:len(exact)] == exact
    # device form with a graph, two replays
    dev = torch.device("cuda:0")
    R = (fa.Request * len(reqs))()
    keep, outs = [], []
    for k, q in enumerate(reqs):
        ta = torch.from_numpy(q["a"].view(np.int64)).to(dev)
        tb = torch.from_numpy(q["b"].view(np.int64)).to(dev)
        tz = torch.zeros((q["zh"] - q["zl"]) * q["zw"], dtype=torch.int64, device=dev)
        keep += [ta, tb, tz]
        outs.append(tz)
        r = R[k]
        r.kind = fa.REQ_FMPZ
        r.a, r.b, r.z = ta.data_ptr(), tb.data_ptr(), tz.data_ptr()
        r.aw, r.bw, r.zw = q["aw"], q["bw"], q["zw"]
        r.an, r.bn = len(q["a"]) // q["aw"], len(q["b"]) // q["bw"]
        r.zl, r.zh = q["zl"], q["zh"]
        r.abits = fa.fmpz_vec_max_bits(q["a"], q["aw"])
        r.bbits = fa.fmpz_vec_max_bits(q["b"], q["bw"])
    B = fa.DeviceBatch(R, graph=True, ks=True)
    stream = torch.cuda.Stream()
    try:
        for rep in range(2):
            for tz in outs:
                tz.zero_()
            torch.cuda.synchronize()
            B.launch(ctypes.c_void_p(stream.cuda_stream))
            stream.synchronize()
            for k in range(len(reqs)):
                assert np.array_equal(outs[k].cpu().numpy().view(np.uint64), got[k][1]), (rep, k)
    finally:
        B.close()
