"""Seeded synthetic request streams of the batched workloads and the runners
that time them: one product per call (the drop-in entry points), one
host-pointer batch, and a device-resident batch replayed from a hipGraph.

    nmod4096   4096 nmod products, lengths log-uniform in [2^6, 2^16], each
               pair with its own random prime modulus of 20..63 bits,
               coefficients uniform in [0, n)
    fmpz256    256 fmpz products, lengths uniform in [1024, 16384], one bit
               size b per pair uniform in [64, 4096], signed coefficients
               uniform in (-2^b, 2^b)
    stream     a mixed stream: 50% integer products (each operand
               log-uniform in [2^4, 2^22] limbs), 30% nmod as nmod4096,
               20% fmpz as fmpz256

fmpz products beyond the eight-prime capacity are computed by Kronecker
substitution onto the integer product in every leg (the route FLINT's
_fmpz_poly_mul takes when fft_small declines).

Sizes are drawn with numpy's seeded generator (seed 0xF1A70000 + the
workload's index: 3, 4, 5).  Limb-products of a request: an*bn for integer
and nmod products, an*bn*wa*wb for fmpz with w = ceil((b+1)/64).
"""
import ctypes
import time

import numpy as np

import flint_amd as fa


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
        if _is_prime(n):
            return n


def signed_coeffs(rng, n, bits, w):
    """n coefficients uniform in (-2^bits, 2^bits) as w-limb two's complement"""
    mag = rng.integers(0, 2**64, (n, w), dtype=np.uint64, endpoint=False)
    full = bits // 64
    mag[:, full + 1:] = 0
    if full < w:
        rem = bits - 64 * full
        mag[:, full] &= np.uint64((1 << rem) - 1) if rem else np.uint64(0)
    neg = rng.integers(0, 2, n).astype(bool)
    out = mag.copy()
    if neg.any():
        v = ~mag[neg]
        carry = np.ones(len(v), dtype=np.uint64)
        for k in range(w):
            s = v[:, k] + carry
            carry = ((s == 0) & (carry == 1)).astype(np.uint64)
            v[:, k] = s
        out[neg] = v
    return out.ravel()


def nmod_request(rng):
    an = int(2 ** rng.uniform(6, 16))
    bn = int(2 ** rng.uniform(6, 16))
    n = random_prime(rng, int(rng.integers(20, 64)))
    a = rng.integers(0, n, an, dtype=np.uint64)
    b = rng.integers(0, n, bn, dtype=np.uint64)
    return {"kind": "nmod", "a": a, "b": b, "n": n, "zl": 0, "zh": an + bn - 1, "lp": an * bn}


def fmpz_request(rng):
    an = int(rng.integers(1024, 16385))
    bn = int(rng.integers(1024, 16385))
    bits = int(rng.integers(64, 4097))
    w = bits // 64 + 1
    zw = (2 * bits + 64) // 64 + 1
    a = signed_coeffs(rng, an, bits, w)
    b = signed_coeffs(rng, bn, bits, w)
    wa = -(-(bits + 1) // 64)
    return {"kind": "fmpz", "a": a, "aw": w, "b": b, "bw": w, "zl": 0, "zh": an + bn - 1,
            "zw": zw, "abits": fa.fmpz_vec_max_bits(a, w), "bbits": fa.fmpz_vec_max_bits(b, w),
            "lp": an * bn * wa * wa}


def mpn_request(rng):
    an = int(2 ** rng.uniform(4, 22))
    bn = int(2 ** rng.uniform(4, 22))
    a = rng.integers(0, 2**64, an, dtype=np.uint64, endpoint=False)
    b = rng.integers(0, 2**64, bn, dtype=np.uint64, endpoint=False)
    a[-1] |= np.uint64(1)
    b[-1] |= np.uint64(1)
    return {"kind": "mpn", "a": a, "b": b, "lp": an * bn}


def make(workload, n=None):
    if workload == "nmod4096":
        rng = np.random.default_rng(0xF1A70003)
        return [nmod_request(rng) for _ in range(n or 4096)]
    if workload == "fmpz256":
        rng = np.random.default_rng(0xF1A70004)
        return [fmpz_request(rng) for _ in range(n or 256)]
    if workload == "stream":
        rng = np.random.default_rng(0xF1A70005)
        out = []
        for _ in range(n or 200000):
            u = rng.random()
            out.append(mpn_request(rng) if u < 0.5 else
                       nmod_request(rng) if u < 0.8 else fmpz_request(rng))
        return out
    raise ValueError(workload)


def plain(reqs):
    """one drop-in call per product (host pointers); (seconds, statuses)"""
    st = []
    t0 = time.perf_counter()
    for q in reqs:
        if q["kind"] == "mpn":
            fa.mpn_mul(q["a"], q["b"])
            st.append(1)
        elif q["kind"] == "nmod":
            fa.nmod_poly_mul_mid(q["a"], q["b"], q["zl"], q["zh"], q["n"])
            st.append(1)
        else:
            z = fa.fmpz_poly_mul_mid(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"], q["zw"])
            if z is None:
                fa.fmpz_poly_mul_ks(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"], q["zw"])
            st.append(1)
    return time.perf_counter() - t0, st


def with_plans(reqs, net):
    """integer requests get the planner's (np, bits)"""
    idx = [k for k, q in enumerate(reqs) if q["kind"] == "mpn"]
    if net is None or not idx:
        return reqs
    shapes = np.array([[len(reqs[k]["a"]), len(reqs[k]["b"])] for k in idx], dtype=np.int64)
    plans = net.plan_mpn(shapes)
    out = list(reqs)
    for k, (np_, bits) in zip(idx, plans):
        q = dict(reqs[k])
        q["np"], q["bits"] = int(np_), int(bits)
        out[k] = q
    return out


def device_batch(reqs, torch, dev, ks=True):
    """uploads the operands and builds a graph-captured device batch;
    returns (batch, outputs, keep-alive list)"""
    R = (fa.Request * len(reqs))()
    keep, outs = [], []
    for k, q in enumerate(reqs):
        r = R[k]
        ta = torch.from_numpy(np.ascontiguousarray(q["a"]).view(np.int64)).to(dev)
        tb = torch.from_numpy(np.ascontiguousarray(q["b"]).view(np.int64)).to(dev)
        kind = q["kind"]
        if kind == "mpn":
            r.kind = fa.REQ_MPN
            r.an, r.bn = len(q["a"]), len(q["b"])
            zwords = r.an + r.bn
            r.np, r.bits = q.get("np", 0), q.get("bits", 0)
        elif kind == "nmod":
            r.kind = fa.REQ_NMOD
            r.an, r.bn = len(q["a"]), len(q["b"])
            r.zl, r.zh = q["zl"], q["zh"]
            r.mod = fa.nmod_init(q["n"])
            zwords = q["zh"] - q["zl"]
        else:
            r.kind = fa.REQ_FMPZ
            r.aw, r.bw, r.zw = q["aw"], q["bw"], q["zw"]
            r.an, r.bn = len(q["a"]) // q["aw"], len(q["b"]) // q["bw"]
            r.zl, r.zh = q["zl"], q["zh"]
            r.abits, r.bbits = q["abits"], q["bbits"]
            zwords = (q["zh"] - q["zl"]) * q["zw"]
        tz = torch.zeros(max(zwords, 1), dtype=torch.int64, device=dev)
        r.a, r.b, r.z = ta.data_ptr(), tb.data_ptr(), tz.data_ptr()
        keep += [ta, tb, tz]
        outs.append(tz)
    B = fa.DeviceBatch(R, graph=True, ks=ks)
    keep.append(R)
    return B, outs, keep


def time_device(B, torch, steps, warmup):
    stream = torch.cuda.Stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    for _ in range(warmup):
        B.launch(sp)
    stream.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        B.launch(sp)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps
