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

Every draw comes from one xoshiro256** stream per workload, seeded by
SplitMix64 with 0xF1A70000 + the workload's config index (3, 4, 5), through
the library's generator (fa_rng_*): uniform integers by rejection,
log-uniform sizes as floor(2^U), primes as odd uniform draws until
Miller-Rabin accepts.  Per request the draws are, in order:
    nmod:  an, bn (log-uniform), modulus bits (uniform), the prime, a, b
    fmpz:  an, bn (uniform), b (uniform), a, b (signed, b // 64 + 1 limbs)
    mpn:   an, bn (log-uniform), a, b (limbs, top limb nonzero)
and in the stream a uniform double in [0, 1) first picks the kind.
Limb-products of a request: an*bn for integer and nmod products,
an*bn*wa*wb for fmpz with w = ceil((b+1)/64).
"""
import ctypes
import time

import numpy as np

import flint_amd as fa

SEEDS = {"nmod4096": 0xF1A70003, "fmpz256": 0xF1A70004, "stream": 0xF1A70005}
COUNTS = {"nmod4096": 4096, "fmpz256": 256, "stream": 200000}


class Rng:
    """the library's seeded generator (fa_rng_*)"""

    def __init__(self, seed):
        self.s = (ctypes.c_uint64 * 4)()
        L = fa.lib()
        self._L = L
        L.fa_rng_seed(self.s, seed)

    def uniform(self, lo, hi):
        return int(self._L.fa_rng_uniform(self.s, lo, hi))

    def double(self):
        return float(self._L.fa_rng_double(self.s))

    def log_uniform(self, x, y):
        return int(self._L.fa_rng_log_uniform(self.s, x, y))

    def prime(self, bits):
        return int(self._L.fa_rng_random_prime(self.s, bits))

    def below(self, n, bound):
        out = np.empty(n, dtype=np.uint64)
        self._L.fa_rng_fill_below(self.s, out.ctypes.data, n, bound)
        return out

    def limbs(self, n):
        out = np.empty(n, dtype=np.uint64)
        self._L.fa_rng_fill_limbs(self.s, out.ctypes.data, n, 1)
        return out

    def signed(self, n, bits, w):
        out = np.empty(n * w, dtype=np.uint64)
        self._L.fa_rng_fill_signed(self.s, out.ctypes.data, n, bits, w)
        return out


def nmod_request(r):
    an = r.log_uniform(6, 16)
    bn = r.log_uniform(6, 16)
    n = r.prime(r.uniform(20, 63))
    a = r.below(an, n)
    b = r.below(bn, n)
    return {"kind": "nmod", "a": a, "b": b, "n": n, "zl": 0, "zh": an + bn - 1, "lp": an * bn}


def fmpz_request(r):
    an = r.uniform(1024, 16384)
    bn = r.uniform(1024, 16384)
    bits = r.uniform(64, 4096)
    w = bits // 64 + 1
    a = r.signed(an, bits, w)
    b = r.signed(bn, bits, w)
    zw = (2 * bits + 64) // 64 + 1
    wa = -(-(bits + 1) // 64)
    return {"kind": "fmpz", "a": a, "aw": w, "b": b, "bw": w, "zl": 0, "zh": an + bn - 1,
            "zw": zw, "bits": bits, "abits": fa.fmpz_vec_max_bits(a, w),
            "bbits": fa.fmpz_vec_max_bits(b, w), "lp": an * bn * wa * wa}


def mpn_request(r):
    an = r.log_uniform(4, 22)
    bn = r.log_uniform(4, 22)
    a = r.limbs(an)
    b = r.limbs(bn)
    return {"kind": "mpn", "a": a, "b": b, "lp": an * bn}


def generate(workload, n=None):
    """yields the requests of a workload in order (n: how many; default all)"""
    if workload not in SEEDS:
        raise ValueError(workload)
    r = Rng(SEEDS[workload])
    for _ in range(COUNTS[workload] if n is None else n):
        if workload == "nmod4096":
            yield nmod_request(r)
        elif workload == "fmpz256":
            yield fmpz_request(r)
        else:
            u = r.double()
            yield mpn_request(r) if u < 0.5 else nmod_request(r) if u < 0.8 else fmpz_request(r)


def make(workload, n=None):
    return list(generate(workload, n))


def res_words(depth):
    """doubles of one prime's residue array (rows of 4096 at a pitch of 4224
    above depth 12)"""
    return (4224 << (depth - 12)) if depth > 12 else (1 << depth)


def mpn_workspace(an, bn):
    p = fa.mpn_plan(an, bn)
    return 16 * p["np"] * res_words(p["depth"]) + 16 * (an + bn)


def request_workspace(q):
    """device workspace of a request (residue rows, Kronecker scratch), for
    cutting a stream into batches"""
    if q["kind"] == "mpn":
        return mpn_workspace(len(q["a"]), len(q["b"]))
    an = len(q["a"]) // q.get("aw", 1)
    bn = len(q["b"]) // q.get("bw", 1)
    depth = max(4, int(np.ceil(np.log2(an + bn))))
    ws = 16 * 8 * res_words(depth) + 8 * (an + bn) * q.get("zw", 1)
    if q["kind"] == "fmpz":
        sb = abs(q["abits"]) + abs(q["bbits"]) + int(min(an, bn)).bit_length() + 1
        la, lb = -(-an * sb // 64), -(-bn * sb // 64)
        ws += mpn_workspace(la, lb) + 8 * 2 * (la + lb)   # one product of the packed operands
    return ws


def chunks(reqs, cap):
    """request indices in consecutive groups of at most `cap` bytes of
    device workspace (a request above the cap is a group of its own)"""
    out, cur, ws = [], [], 0
    for k, q in enumerate(reqs):
        w = request_workspace(q)
        if cur and ws + w > cap:
            out.append(cur)
            cur, ws = [], 0
        cur.append(k)
        ws += w
    if cur:
        out.append(cur)
    return out


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
    for k, (np_, bits, tr) in zip(idx, plans):
        q = dict(reqs[k])
        q["np"], q["bits"] = int(np_), int(bits)
        q["trunc"] = fa.TRUNC_ON if tr else fa.TRUNC_OFF
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
            r.trunc = q.get("trunc", fa.TRUNC_AUTO)
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
    # the uploads and fills ran on torch's stream; the batch runs on others
    torch.cuda.synchronize()
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


def device_per_call(reqs, R, stream_ptr):
    """one device entry-point call per request of a recorded batch's request
    array R (device operands and outputs already in place), enqueued on one
    stream: the per-call counterpart of the batch replay"""
    L = fa.lib()
    for q, r in zip(reqs, R):
        if q["kind"] == "mpn":
            rc = L.fa_mpn_mul_device(r.z, r.a, r.an, r.b, r.bn, stream_ptr)
        elif q["kind"] == "nmod":
            rc = L.fa_nmod_poly_mul_mid_device(r.z, r.zl, r.zh, r.a, r.an, r.b, r.bn, r.mod,
                                               stream_ptr)
        else:
            rc = L.fa_fmpz_poly_mul_mid_device(r.z, r.zw, r.zl, r.zh, r.a, r.aw, r.an, r.b, r.bw,
                                               r.bn, r.abits, r.bbits, 1, stream_ptr)
        if rc < 0:
            raise fa.FlintAmdError("device call failed: %s" % L.fa_last_error().decode())


def time_device_per_call(reqs, R, torch, steps, warmup):
    """ms per pass of device_per_call, HIP events on its stream"""
    stream = torch.cuda.Stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    for _ in range(warmup):
        device_per_call(reqs, R, sp)
    stream.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        device_per_call(reqs, R, sp)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


# ----------------------------------------------------------------------
# the mixed stream at its full size (200,000 requests, about 0.5 TB of
# operands), generated, multiplied and checked on the device chunk by chunk
# ----------------------------------------------------------------------

# 61-bit prime of the integer and fmpz checks
CHECK_Q = 2305843009213693951


def stream_words(h):
    """(a words, b words, z words, a width, z width) of a stream header"""
    an, bn = h.an, h.bn
    if h.kind == fa.REQ_MPN:
        return an, bn, an + bn, 1, 1
    if h.kind == fa.REQ_NMOD:
        return an, bn, an + bn - 1, 1, 1
    w = h.bits // 64 + 1
    zw = (2 * h.bits + 64) // 64 + 1
    return an * w, bn * w, (an + bn - 1) * zw, w, zw


def stream_header_workspace(h):
    a, b, z, w, zw = stream_words(h)
    io = 8 * (a + b + z)
    if h.kind == fa.REQ_MPN:
        return io + mpn_workspace(h.an, h.bn)
    depth = max(4, int(np.ceil(np.log2(h.an + h.bn))))
    ws = io + 16 * 8 * res_words(depth)
    if h.kind == fa.REQ_FMPZ:
        sb = 2 * h.bits + int(min(h.an, h.bn)).bit_length() + 1
        la, lb = -(-h.an * sb // 64), -(-h.bn * sb // 64)
        ws += mpn_workspace(la, lb) + 8 * 2 * (la + lb)
    return ws


def stream_chunks(H, cap):
    """header indices in consecutive groups of at most `cap` bytes of
    operands, outputs and workspace"""
    out, cur, ws = [], [], 0
    for k in range(len(H)):
        w = stream_header_workspace(H[k])
        if cur and ws + w > cap:
            out.append(cur)
            cur, ws = [], 0
        cur.append(k)
        ws += w
    if cur:
        out.append(cur)
    return out


def stream_limb_products(h):
    if h.kind == fa.REQ_FMPZ:
        wa = -(-(h.bits + 1) // 64)
        return h.an * h.bn * wa * wa
    return h.an * h.bn


class StreamChunk:
    """one chunk of the full stream in HBM: operands written on the device,
    requests recorded into a graph-captured batch, outputs checked by
    evaluation"""

    def __init__(self, H, idx, torch, dev, net=None):
        self.H, self.idx, self.torch = H, idx, torch
        sizes = [stream_words(H[k]) for k in idx]
        total = sum(a + b + z for a, b, z, _, _ in sizes)
        self.arena = torch.empty(max(total, 1), dtype=torch.int64, device=dev)
        base = self.arena.data_ptr()
        R = (fa.Request * len(idx))()
        F = (fa.FillSpec * (2 * len(idx)))()
        off = 0
        self.ptrs = []
        for j, (k, (aw_, bw_, zw_, w, zw)) in enumerate(zip(idx, sizes)):
            h = H[k]
            pa, pb, pz = base + 8 * off, base + 8 * (off + aw_), base + 8 * (off + aw_ + bw_)
            off += aw_ + bw_ + zw_
            self.ptrs.append((pa, pb, pz))
            for s_, (p, cnt) in enumerate(((pa, h.an), (pb, h.bn))):
                f = F[2 * j + s_]
                f.dst, f.count, f.key, f.kind = p, cnt, h.key[s_], h.kind
                f.n, f.w, f.bits = h.n, w, h.bits
            r = R[j]
            r.kind, r.an, r.bn = h.kind, h.an, h.bn
            r.a, r.b, r.z = pa, pb, pz
            if h.kind == fa.REQ_NMOD:
                r.zl, r.zh = 0, h.an + h.bn - 1
                r.mod = fa.nmod_init(h.n)
            elif h.kind == fa.REQ_FMPZ:
                r.aw = r.bw = w
                r.zw = zw
                r.zl, r.zh = 0, h.an + h.bn - 1
                # the generator's bound: |c| < 2^b with both signs present
                r.abits = r.bbits = -h.bits
        if net is not None:
            mi = [j for j, k in enumerate(idx) if H[k].kind == fa.REQ_MPN]
            if mi:
                shapes = np.array([[H[idx[j]].an, H[idx[j]].bn] for j in mi], dtype=np.int64)
                for j, (np_, bits, tr) in zip(mi, net.plan_mpn(shapes)):
                    R[j].np, R[j].bits = int(np_), int(bits)
                    R[j].trunc = fa.TRUNC_ON if tr else fa.TRUNC_OFF
        self.R, self.F = R, F
        torch.cuda.synchronize()            # the arena comes from torch's stream
        fa.stream_fill_device(F)
        self.batch = fa.DeviceBatch(R, graph=True, ks=True)
        self.limb_products = sum(stream_limb_products(H[k]) for k in idx)

    def run(self, stream):
        self.batch.launch(stream)

    def check(self):
        """the requests whose product fails A(x) B(x) = Z(x) mod m"""
        H, idx = self.H, self.idx
        E = (fa.EvalSpec * (3 * len(idx)))()
        mods = []
        for j, k in enumerate(idx):
            h = H[k]
            pa, pb, pz = self.ptrs[j]
            _, _, _, w, zw = stream_words(h)
            if h.kind == fa.REQ_MPN:
                m, x, sg, cnt = CHECK_Q, (1 << 64) % CHECK_Q, 0, (h.an, h.bn, h.an + h.bn)
                ws = (1, 1, 1)
            elif h.kind == fa.REQ_NMOD:
                m, x, sg, cnt = h.n, h.key[0] % h.n, 0, (h.an, h.bn, h.an + h.bn - 1)
                ws = (1, 1, 1)
            else:
                m, x, sg, cnt = CHECK_Q, h.key[1] % CHECK_Q, 1, (h.an, h.bn, h.an + h.bn - 1)
                ws = (w, w, zw)
            mods.append(m)
            for t, p in enumerate((pa, pb, pz)):
                e = E[3 * j + t]
                e.src, e.count, e.w, e.x, e.m, e.is_signed = p, cnt[t], ws[t], x, m, sg
        v = fa.eval_device(E)
        bad = []
        for j, k in enumerate(idx):
            a, b, z = (int(t) for t in v[3 * j:3 * j + 3])
            if a * b % mods[j] != z or self.R[j].status != 1:
                bad.append(k)
        return bad

    def close(self):
        self.batch.close()
        self.arena = None


def run_full_stream(torch, dev, count=None, cap=48 << 30, net=None, per_call=False,
                    progress=None, on_chunk=None):
    """every request of the full stream, chunk by chunk: batch replay timed
    with HIP events (one untimed replay first), products checked; optionally
    the same requests one device call each; on_chunk(chunk) runs before a
    chunk is released.  Returns a summary dict."""
    H = fa.stream_headers(COUNTS["stream"] if count is None else count)
    groups = stream_chunks(H, cap)
    stream = torch.cuda.Stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    tot = {"requests": len(H), "chunks": len(groups), "limb_products": 0, "batch_ms": 0.0,
           "per_call_ms": 0.0 if per_call else None, "failed": [], "wall_s": 0.0}
    t0 = time.perf_counter()
    for c, idx in enumerate(groups):
        ch = StreamChunk(H, idx, torch, dev, net)
        ch.run(sp)
        stream.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ch.run(sp)
        e1.record(stream)
        stream.synchronize()
        ms = e0.elapsed_time(e1)
        bad = ch.check()
        pc = None
        if per_call:
            reqs = [{"kind": {0: "mpn", 1: "nmod", 2: "fmpz"}[H[k].kind]} for k in idx]
            e0.record(stream)
            device_per_call(reqs, ch.R, sp)
            e1.record(stream)
            stream.synchronize()
            pc = e0.elapsed_time(e1)
            tot["per_call_ms"] += pc
            bad = sorted(set(bad) | set(ch.check()))
        if on_chunk:
            on_chunk(ch)
        tot["limb_products"] += ch.limb_products
        tot["batch_ms"] += ms
        tot["failed"] += bad
        ch.close()
        del ch
        torch.cuda.empty_cache()            # chunk arenas differ in size
        if progress:
            progress("chunk %d/%d: %d requests, batch %.1f ms%s, failed %d" %
                     (c + 1, len(groups), len(idx), ms,
                      "" if pc is None else ", per call %.1f ms" % pc, len(bad)))
    tot["wall_s"] = time.perf_counter() - t0
    return tot
