#!/usr/bin/env python3
"""bench.py -- device-resident throughput of the MI355X fft_small path.

Workload (N=1): plain mpn_mul of two naturals of 2^24 64-bit limbs each,
uniform random limbs from the seeded generator (seed 0xF1A70002, top limb
nonzero), one product per call, operands and product resident in HBM.
A "step" is one full product.  The metric is limb-products/s = an*bn per
second.  By default 30 untimed products come first: back-to-back products
of this size take about 25 ms to reach their steady rate.  After the timed region the product is checked against
(a mod q)(b mod q) for three 61-bit primes q.

The roofline entry is for the dominant kernel: algorithmic bytes per launch
(8 bytes per transform point per array pass, at the reference plan's
geometry np*T) divided by that kernel's average launch time, measured with
HIP events recorded around each launch on the launching stream.  The timed
products run the primes in two groups on forked streams (--prime-groups);
the per-kernel table is taken with one launch per pass over all primes.

The CPU baseline describes the host (CPU model, core count), times FLINT's
own mpn_mul_default_mpn_ctx at 1 and nproc threads when a system libflint is
present, and otherwise GMP's mpn_mul and the oracle (oracle/fsmall_oracle.c,
a scalar C restatement of the reference) on the same inputs, each checked
word for word against the GPU product; FLINT's own times measured on
another host are carried beside them, labelled as such.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
FP64_LANE_OPS_PEAK = 3.93e13  # 78.6 TFLOP/s FP64 vector = 3.93e13 lane-ops/s

# reference-plan geometry (mpn_ctx_best_profile at 1 thread, mpn_mul.c:1248-1340)
WORKLOADS = {
    "mpn_2p24": dict(key="mpn_2p24", an=1 << 24, bn=1 << 24, seed=0xF1A70002, ref_np=6,
                     ref_bits=136,
                     ref_T=15790336, ref_depth=24,
                     desc="plain mpn_mul, 2^24 x 2^24 limbs, one product per call"),
    "mpn_65536": dict(key="mpn_65536", an=65536, bn=65536, seed=0xF1A70001, ref_np=6,
                      ref_bits=140,
                      ref_T=60160, ref_depth=16,
                      desc="plain mpn_mul, 65536 x 65536 limbs, one product per call"),
}

# array passes (8 bytes per point each) of every kernel of the mpn pipeline
KERNEL_PASSES = {
    "fa_k_col_fwd16w": 2,  # writes the transformed a and b columns; reads the limbs
    "fa_k_col_inv16w": 2,  # reads and writes the product columns
    "fa_k_col_fwd16": 2,
    "fa_k_row_conv16": 3,  # reads a and b rows, writes the product row
    "fa_k_col_inv16": 2,   # reads and writes the product columns
    "fa_k_crt_fp": 1,      # reads the residues; writes the output limbs
    "fa_k_crt_mpn16": 1,
    "fa_k_col_fwd": 2,
    "fa_k_row_conv": 3,
    "fa_k_col_inv": 2,
    "fa_k_crt_mpn": 1,
    "fa_k_small_conv": 1,
}


def kernel_bytes(name, W):
    pts = W["ref_np"] * W["ref_T"]
    b = 8 * KERNEL_PASSES.get(name, 0) * pts
    if name.startswith("fa_k_col_fwd"):
        b += 8 * (W["an"] + W["bn"])
    if name.startswith("fa_k_crt_mpn") or name == "fa_k_crt_fp":
        b += 8 * (W["an"] + W["bn"])
    return b


def product_model(W):
    depth = W["ref_depth"]
    k = 0 if depth <= 12 else 6 * ((depth + 11) // 12) - 2
    B = 8 * (W["an"] + W["bn"]) * 2 + 8 * W["ref_np"] * W["ref_T"] * k
    F = W["ref_np"] * W["ref_T"] * (13.5 * depth + 12)
    return B, F


def mod_q(limbs, q):
    return int.from_bytes(np.ascontiguousarray(limbs).tobytes(), "little") % q


# FP64 VALU lane-ops per transform point of each kernel of the two-level
# pipeline (4.5 per point per radix-2 level: the radix-4 count of
# sd_fft.c:331-355; 12 per point for the pointwise product with the
# normalizer): the column passes do 12 levels of each operand (forward) or of
# the product (inverse), the row pass 12 forward levels of both operands, the
# pointwise product and 12 inverse levels
KERNEL_FP64_OPS = {
    "fa_k_col_fwd16w": 2 * 12 * 4.5,
    "fa_k_col_fwd16": 2 * 12 * 4.5,
    "fa_k_row_conv16": 2 * 12 * 4.5 + 12 + 12 * 4.5,
    "fa_k_col_inv16w": 12 * 4.5,
    "fa_k_col_inv16": 12 * 4.5,
}

# FLINT's own CPU path on the same configurations, measured on another host
# (Intel Xeon, Sapphire Rapids flags, 8 vCPU; FLINT 3.7.0-dev,
# GMP 6.2.1): not this host
FLINT_OTHER_HOST = {
    "mpn_2p24": {"1_thread_s": 3.13, "8_threads_s": 0.694},
    "mpn_65536": {"1_thread_s": 0.01010, "8_threads_s": 0.00210},
}


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def find_libflint():
    """a system libflint exporting mpn_mul_default_mpn_ctx, or None"""
    import ctypes
    import subprocess
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True,
                             timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    for line in out.splitlines():
        if "libflint" in line and "=>" in line:
            path = line.split("=>", 1)[1].strip()
            try:
                L = ctypes.CDLL(path)
                if hasattr(L, "mpn_mul_default_mpn_ctx"):
                    return path
            except OSError:
                continue
    return None


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(W, a_h, b_h, z_gpu, budget_s=30.0):
    """The CPU legs: host description; FLINT's CPU path if a system
    libflint is present (1 and nproc threads); otherwise GMP's mpn_mul and the
    oracle (the scalar C restatement), labelled, on the same inputs; and
    FLINT's own times from another host, labelled as such"""
    import ctypes
    info = host_info()
    an, bn = W["an"], W["bn"]
    lp = an * bn
    res = {"unit": "limb-products/s", "host": info}
    legs = {}
    flint = find_libflint()
    res["system_libflint"] = flint
    if flint:
        L = ctypes.CDLL(flint)
        L.mpn_mul_default_mpn_ctx.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                              ctypes.c_void_p, ctypes.c_long]
        L.flint_set_num_threads.argtypes = [ctypes.c_int]
        z = np.empty(an + bn, dtype=np.uint64)
        for th in (1, info["usable_cpus"]):
            L.flint_set_num_threads(th)
            L.mpn_mul_default_mpn_ctx(z.ctypes.data, a_h.ctypes.data, an, b_h.ctypes.data, bn)
            t = _median_time(lambda: L.mpn_mul_default_mpn_ctx(
                z.ctypes.data, a_h.ctypes.data, an, b_h.ctypes.data, bn), 5)
            legs["flint_%d_threads" % th] = {"s": t, "value": lp / t, "threads": th,
                                             "bit_exact_vs_gpu": bool(np.array_equal(z, z_gpu))}
    # GMP mpn_mul (libgmp.so.10), the oracle of the reference's t-mul.c:54
    try:
        G = ctypes.CDLL("libgmp.so.10")
        G.__gmpn_mul.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                 ctypes.c_void_p, ctypes.c_long]
        z = np.empty(an + bn, dtype=np.uint64)
        x, y = (a_h, b_h) if an >= bn else (b_h, a_h)
        t0 = time.perf_counter()
        G.__gmpn_mul(z.ctypes.data, x.ctypes.data, len(x), y.ctypes.data, len(y))
        t = time.perf_counter() - t0
        legs["gmp_mpn_mul_1_thread"] = {"s": t, "value": lp / t, "threads": 1,
                                        "bit_exact_vs_gpu": bool(np.array_equal(z, z_gpu)),
                                        "library": "libgmp.so.10 __gmpn_mul, full product"}
    except OSError as e:
        legs["gmp_mpn_mul_1_thread"] = {"unavailable": str(e)}
    # the oracle, full product when it fits the budget, else a sample scaled
    # by n log n
    from oracle import oracle as O
    n = 1 << 16
    rng = np.random.default_rng(12345)
    sa = rng.integers(0, 2**64, n, dtype=np.uint64)
    sb = rng.integers(0, 2**64, n, dtype=np.uint64)
    sa[-1] |= np.uint64(1)
    sb[-1] |= np.uint64(1)
    t0 = time.perf_counter()
    O.mpn_mul(sa, sb)
    t_small = time.perf_counter() - t0
    N = max(an, bn)
    scale = (N * np.log2(N)) / (n * np.log2(n))
    if t_small * scale <= budget_s:
        t0 = time.perf_counter()
        zo = O.mpn_mul(a_h, b_h)
        t_or = time.perf_counter() - t0
        legs["oracle_1_thread"] = {"s": t_or, "value": lp / t_or, "threads": 1,
                                   "bit_exact_vs_gpu": bool(np.array_equal(zo, z_gpu)),
                                   "sample": "full product"}
    else:
        t_or = t_small * scale
        legs["oracle_1_thread"] = {"s": t_or, "value": lp / t_or, "threads": 1,
                                   "sample": "2^16 x 2^16 took %.3f s, scaled by n log n (x%.1f)"
                                             % (t_small, scale)}
    ref = FLINT_OTHER_HOST.get(W.get("key"))
    if ref:
        res["flint_same_code_different_host"] = {
            k: {"s": v, "value": lp / v} for k, v in ref.items()}
        res["flint_same_code_different_host"]["host"] = (
            "Intel Xeon (Sapphire Rapids flags), 8 vCPU, FLINT 3.7.0-dev, "
            "GMP 6.2.1")
    res["legs"] = legs
    if flint:
        best = max((v for k, v in legs.items() if k.startswith("flint")), key=lambda v: v["value"])
        res.update(value=best["value"], cores=best["threads"], kind="reference",
                   sample="system libflint mpn_mul_default_mpn_ctx, full product, median of 5")
    else:
        o = legs["oracle_1_thread"]
        res.update(value=o["value"], cores=1, kind="port",
                   sample="the oracle (scalar C restatement of the reference), %s, 1 thread; "
                          "no system libflint on this host; GMP mpn_mul beside it in legs"
                          % o.get("sample", ""))
    return res


def _k_passes(depth):
    """array passes of the byte model at a transform depth (0: one LDS-resident
    kernel; 10 for a two-level transform)"""
    return 0 if depth <= 12 else 6 * ((depth + 11) // 12) - 2


def _mpn_model_bytes(an, bn, np_, bits, depth):
    T = -(-64 * an // bits) + -(-64 * bn // bits) - 1
    return 16 * (an + bn) + 8 * np_ * T * _k_passes(depth)


def request_model_bytes(q, r, fa):
    """algorithmic HBM bytes of one request at the plan the batch chose
    (np, bits, depth written back into the request r)"""
    if q["kind"] == "mpn":
        return _mpn_model_bytes(len(q["a"]), len(q["b"]), r.np, r.bits, r.depth)
    aw, bw = q.get("aw", 1), q.get("bw", 1)
    an, bn = len(q["a"]) // aw, len(q["b"]) // bw
    zn = an + bn - 1
    zw = q.get("zw", 1)
    io = 8 * (an * aw + bn * bw + (q["zh"] - q["zl"]) * zw)
    if r.np > 0:
        return io + 8 * r.np * zn * _k_passes(r.depth)
    # fmpz beyond eight primes: one integer product of the signed-slot
    # packed operands (packed limbs written, read by the transform, product
    # limbs read back by the unpack)
    sb = abs(q["abits"]) + abs(q["bbits"]) + int(min(an, bn)).bit_length() + 1
    la, lb = -(-an * sb // 64), -(-bn * sb // 64)
    p = fa.mpn_plan(la, lb)
    return io + 16 * (la + lb) + _mpn_model_bytes(la, lb, p["np"], p["bits"], p["depth"])


def batched_main(args, torch, fa):
    """batched workloads: one drop-in call per product, one host batch, and a
    device-resident batch replayed from a hipGraph, on the same requests"""
    from flint_amd import workloads as WL
    from flint_amd import planner as PL
    t0 = time.time()
    reqs = WL.make(args.workload, args.stream_n if args.workload == "stream" else None)
    print("# %d requests generated in %.1fs" % (len(reqs), time.time() - t0), file=sys.stderr,
          flush=True)
    net = None
    if not args.no_planner and os.path.exists(PL.WEIGHTS):
        net = PL.load_planner()
    reqs_p = WL.with_plans(reqs, net)
    lp = sum(q["lp"] for q in reqs)

    # one product per call, results kept as the reference for the batches
    plain_out = []
    tp0 = time.perf_counter()
    for q in reqs:
        if q["kind"] == "mpn":
            plain_out.append(fa.mpn_mul(q["a"], q["b"]))
        elif q["kind"] == "nmod":
            plain_out.append(fa.nmod_poly_mul_mid(q["a"], q["b"], q["zl"], q["zh"], q["n"]))
        else:
            z = fa.fmpz_poly_mul_mid(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"], q["zw"])
            if z is None:   # beyond eight primes: Kronecker substitution, as _fmpz_poly_mul
                z = fa.fmpz_poly_mul_ks(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"],
                                        q["zw"])
            plain_out.append(z)
    t_plain = time.perf_counter() - tp0
    done = [k for k, z in enumerate(plain_out) if z is not None]
    lp_done = sum(reqs[k]["lp"] for k in done)
    print("# plain: %.3f s" % t_plain, file=sys.stderr, flush=True)

    # chunks of at most 32 GB of residue workspace
    chunks = WL.chunks(reqs_p, 32 << 30)

    # host-pointer batches: one untimed pass over every chunk sizes the
    # per-thread device and pinned buffers the one-shot batch borrows (to the
    # largest chunk), then the timed pass
    for ch in chunks:
        fa.batch_mul([reqs_p[k] for k in ch], ks=True)
    th0 = time.perf_counter()
    batch_out = {}
    for ch in chunks:
        res = fa.batch_mul([reqs_p[k] for k in ch], ks=True)
        for k, (st, z) in zip(ch, res):
            batch_out[k] = (st, z)
    t_host = time.perf_counter() - th0
    exact = all((batch_out[k][0] == 1) == (plain_out[k] is not None) and
                (plain_out[k] is None or np.array_equal(batch_out[k][1], plain_out[k]))
                for k in range(len(reqs)))
    print("# batched host: %.3f s, bit-exact vs one-per-call: %s" % (t_host, exact),
          file=sys.stderr, flush=True)

    # device-resident batches, graph replays
    dev = torch.device("cuda:0")
    ms_dev = 0.0
    dev_exact = True
    model_bytes = 0
    ms_dev_plain = 0.0
    for ch in chunks:
        sub = [reqs_p[k] for k in ch]
        B, outs, keep = WL.device_batch(sub, torch, dev)
        R = keep[-1]
        model_bytes += sum(request_model_bytes(reqs_p[k], R[j], fa) for j, k in enumerate(ch))
        ms_dev += WL.time_device(B, torch, args.steps, args.warmup)
        for k, tz in zip(ch, outs):
            if plain_out[k] is not None:
                got = tz.cpu().numpy().view(np.uint64)[:len(plain_out[k])]
                dev_exact &= bool(np.array_equal(got, plain_out[k]))
            tz.zero_()
        # the same requests one device call each, same buffers and plans
        # (integer requests on the automatic plan of fa_mpn_mul_device)
        ms_dev_plain += WL.time_device_per_call(sub, R, torch, max(1, args.steps // 2), 1)
        for k, tz in zip(ch, outs):
            if plain_out[k] is not None:
                got = tz.cpu().numpy().view(np.uint64)[:len(plain_out[k])]
                dev_exact &= bool(np.array_equal(got, plain_out[k]))
        B.close()
        del outs, keep
        torch.cuda.empty_cache()
    print("# batched device: %.3f ms per pass; one device call per product: %.3f ms; "
          "bit-exact: %s" % (ms_dev, ms_dev_plain, dev_exact), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        info = host_info()
        nth = min(info["usable_cpus"] or 1, 16)

        def cpu_one(k):
            q = reqs[k]
            if q["kind"] == "mpn":
                O.mpn_mul(q["a"], q["b"])
            elif q["kind"] == "nmod":
                O.nmod_poly_mul_mid(q["a"], q["b"], q["zl"], q["zh"], q["n"])
            else:
                ok, _ = O.fmpz_poly_mul_mid(q["a"], q["aw"], q["b"], q["bw"], q["zl"], q["zh"],
                                            q["zw"])
                if not ok:      # the oracle has no Kronecker route: not part of the sample
                    return 0
            return q["lp"]

        # one request per thread over the usable cores (the oracle releases the
        # GIL), in waves until about 15 s have passed
        tc0 = time.perf_counter()
        lp_s, n_s, pos = 0, 0, 0
        with ThreadPoolExecutor(nth) as ex:
            while pos < len(done) and time.perf_counter() - tc0 < 15.0:
                wave = done[pos:pos + 4 * nth]
                pos += len(wave)
                got = list(ex.map(cpu_one, wave))
                lp_s += sum(got)
                n_s += sum(1 for g in got if g)
        tc = time.perf_counter() - tc0
        cpu = {"value": lp_s / tc, "unit": "limb-products/s", "cores": nth, "kind": "port",
               "host": info, "system_libflint": find_libflint(),
               "sample": "oracle (scalar C restatement) on the first %d computed requests of the "
                         "same stream, one request per thread over %d threads, %.1f s"
                         % (n_s, nth, tc)}

    value = lp_done / (ms_dev * 1e-3)
    out = {
        "metric": "limb-products/s batched vs plain GPU vs CPU, bit-exact; % HBM roofline",
        "value": value, "unit": "limb-products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_dev, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": {"nmod4096": "batched nmod_poly_mul, 4096 products",
                         "fmpz256": "batched fmpz_poly_mul, 256 products",
                         "stream": "mixed request stream sample"}[args.workload],
            "requests": len(reqs), "computed": len(done), "declined": len(reqs) - len(done),
            "chunks": len(chunks), "planner": "learned" if net is not None else "automatic",
            "limb_products": lp_done,
            "plain_one_call_per_product_s": t_plain,
            "plain_limb_products_per_s": lp_done / t_plain,
            "batched_host_s": t_host,
            "batched_host_limb_products_per_s": lp_done / t_host,
            "batched_device_ms": ms_dev,
            "plain_device_one_call_per_product_ms": ms_dev_plain,
            "speedup_batched_device_vs_plain_device": ms_dev_plain / ms_dev,
            "speedup_batched_host_vs_plain_host": t_plain / t_host,
            "speedup_batched_device_vs_plain_host": t_plain / (ms_dev * 1e-3),
            "bit_exact_vs_one_call_per_product": bool(exact and dev_exact),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "whole batch: algorithmic bytes of every request at the plan it ran "
                      "(8 B per transform point per array pass, 10 passes above depth 12, "
                      "operands and outputs once)",
            "achieved": model_bytes / (ms_dev * 1e-3) / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": model_bytes / (ms_dev * 1e-3) / HBM_PEAK,
            "traffic": None,
            "model_bytes": model_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def stream_model_bytes(h, r, fa):
    """algorithmic HBM bytes of one full-stream request at the plan it ran"""
    from flint_amd import workloads as WL
    a_w, b_w, z_w, w, zw = WL.stream_words(h)
    if h.kind == fa.REQ_MPN:
        return _mpn_model_bytes(h.an, h.bn, r.np, r.bits, r.depth)
    io = 8 * (a_w + b_w + z_w)
    zn = h.an + h.bn - 1
    if r.np > 0:
        return io + 8 * r.np * zn * _k_passes(r.depth)
    sb = 2 * h.bits + int(min(h.an, h.bn)).bit_length() + 1
    la, lb = -(-h.an * sb // 64), -(-h.bn * sb // 64)
    p = fa.mpn_plan(la, lb)
    return io + 16 * (la + lb) + _mpn_model_bytes(la, lb, p["np"], p["bits"], p["depth"])


def full_stream_main(args, torch, fa):
    """the mixed stream at its full size (200,000 requests): operands written
    on the device chunk by chunk, each chunk one graph-captured batch (timed
    replay after an untimed one), then the same requests one device call
    each; every product checked by evaluation mod a 61-bit prime (integers,
    fmpz) or mod n (nmod)"""
    from flint_amd import workloads as WL
    from flint_amd import planner as PL
    net = None
    if not args.no_planner and os.path.exists(PL.WEIGHTS):
        net = PL.load_planner()
    dev = torch.device("cuda:0")
    model = [0]

    def on_chunk(ch):
        for j, k in enumerate(ch.idx):
            model[0] += stream_model_bytes(ch.H[k], ch.R[j], fa)

    def progress(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    count = args.stream_n if args.stream_n > 1000 else None
    tot = WL.run_full_stream(torch, dev, count=count, net=net, per_call=True, progress=progress,
                             on_chunk=on_chunk)
    ms = tot["batch_ms"]
    cpu = None
    if not args.no_cpu_baseline:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        info = host_info()
        nth = min(info["usable_cpus"] or 1, 16)
        H = fa.stream_headers(2000)

        def cpu_one(k):
            h = H[k]
            a_w, b_w, z_w, w, zw = WL.stream_words(h)
            a = O.stream_operand(h.kind, h.key[0], h.an, n=h.n, w=w, bits=h.bits)
            b = O.stream_operand(h.kind, h.key[1], h.bn, n=h.n, w=w, bits=h.bits)
            t0 = time.perf_counter()
            if h.kind == 0:
                O.mpn_mul(a, b)
            elif h.kind == 1:
                O.nmod_poly_mul_mid(a, b, 0, h.an + h.bn - 1, h.n)
            else:
                ok, _ = O.fmpz_poly_mul_mid(a, w, b, w, 0, h.an + h.bn - 1, zw)
                if not ok:
                    return 0, 0.0
            return WL.stream_limb_products(h), time.perf_counter() - t0

        tc0 = time.perf_counter()
        lp_s, n_s, pos, busy = 0, 0, 0, 0.0
        with ThreadPoolExecutor(nth) as ex:
            while pos < len(H) and time.perf_counter() - tc0 < 15.0:
                wave = list(range(pos, min(pos + 4 * nth, len(H))))
                pos += len(wave)
                for lp, t in ex.map(cpu_one, wave):
                    lp_s += lp
                    n_s += 1 if lp else 0
                    busy += t
        tc = time.perf_counter() - tc0
        cpu = {"value": lp_s / tc, "unit": "limb-products/s", "cores": nth, "kind": "port",
               "host": info, "system_libflint": find_libflint(),
               "sample": "oracle (scalar C restatement) on the first %d computed requests of the "
                         "stream (operands from the oracle's restatement of the generator), one "
                         "request per thread over %d threads, %.1f s" % (n_s, nth, tc)}
    out = {
        "metric": "limb-products/s batched vs plain GPU vs CPU, bit-exact; % HBM roofline",
        "value": tot["limb_products"] / (ms * 1e-3), "unit": "limb-products/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "mixed request stream, full size (200,000 requests; operands generated "
                        "on the device, every product checked by evaluation)",
            "requests": tot["requests"], "chunks": tot["chunks"],
            "planner": "learned" if net is not None else "automatic",
            "limb_products": tot["limb_products"],
            "batched_device_ms": ms,
            "plain_device_one_call_per_product_ms": tot["per_call_ms"],
            "speedup_batched_device_vs_plain_device": tot["per_call_ms"] / ms,
            "failed_checks": len(tot["failed"]),
            "wall_s_generation_batches_checks": tot["wall_s"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "whole stream: algorithmic bytes of every request at the plan it ran",
            "achieved": model[0] / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": model[0] / (ms * 1e-3) / HBM_PEAK, "traffic": None, "model_bytes": model[0],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 30 for the plain products, 10 for batches)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 30 for the plain products, 10 for "
                         "batches): the first 20-30 ms of back-to-back work run about 5%% "
                         "slower than the steady state")
    ap.add_argument("--workload", default="mpn_2p24",
                    choices=sorted(WORKLOADS) + ["nmod4096", "fmpz256", "stream",
                                                 "stream-full"])
    ap.add_argument("--stream-n", type=int, default=1000,
                    help="requests of the mixed stream sample (workload stream); with "
                         "stream-full, values above 1000 cut the full stream to its first N")
    ap.add_argument("--no-planner", action="store_true",
                    help="batched workloads: automatic plans instead of the learned planner")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--prime-groups", type=int, default=2,
                    help="depth-24 products: primes in this many groups on forked streams "
                         "(0 = one launch per pass, as profiler kernel traces should see them)")
    ap.add_argument("--no-truncation", action="store_true",
                    help="run every transform at its full 2^depth points (outputs identical)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic_mpn_2p24.json"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's plain copy kernel over MIB MiB, 5 times (the "
                         "known-size reference for scaling HBM counter readings), and exit")
    args = ap.parse_args()
    plain = args.workload in WORKLOADS
    if args.steps is None:
        args.steps = 30 if plain else 10
    if args.warmup is None:
        args.warmup = 30 if plain else 10

    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (sharding the primes "
              "or the requests across GPUs is designed but not built)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import flint_amd as fa

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.int64, device="cuda:0")
        dst = torch.empty_like(src)
        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        fa.copy_device(dst.data_ptr(), src.data_ptr(), nbytes, stream.cuda_stream)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(5):
            fa.copy_device(dst.data_ptr(), src.data_ptr(), nbytes, stream.cuda_stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / 5
        assert bool(torch.equal(src, dst))
        print(json.dumps({"copy_bytes": nbytes, "read_plus_write_bytes": 2 * nbytes,
                          "ms": ms, "GB/s": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return

    fa.set_prime_groups(args.prime_groups)
    fa.set_truncation(not args.no_truncation)
    if args.workload == "stream-full":
        full_stream_main(args, torch, fa)
        return
    if args.workload in ("nmod4096", "fmpz256", "stream"):
        return batched_main(args, torch, fa)

    W = WORKLOADS[args.workload]
    an, bn = W["an"], W["bn"]
    t0 = time.time()
    a_h, b_h = fa.synth_mpn_pair(W["seed"], an, bn)
    print("# inputs generated in %.2fs" % (time.time() - t0), file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    a = torch.from_numpy(a_h.view(np.int64)).to(dev)
    b = torch.from_numpy(b_h.view(np.int64)).to(dev)
    z = torch.empty(an + bn, dtype=torch.int64, device=dev)
    # a dedicated (default-priority) stream: the library enqueues on it and the
    # timing events are recorded on it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    plan = fa.mpn_plan(an, bn)

    def step():
        fa.mpn_mul_device(z.data_ptr(), a.data_ptr(), an, b.data_ptr(), bn, sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("# warmup done", file=sys.stderr, flush=True)

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    print("# timed: %.3f ms/product" % ms, file=sys.stderr, flush=True)

    # per-kernel launch times (HIP events around each launch, same stream)
    fa.prof_reset()
    fa.prof_enable(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    fa.prof_enable(False)
    kt = fa.prof_read()
    per_kernel = {k: dict(avg_ms=v[0] / v[1], launches=v[1]) for k, v in kt.items()}
    dominant = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"] * per_kernel[k]["launches"])
    dk = per_kernel[dominant]
    dbytes = kernel_bytes(dominant, W)
    achieved = dbytes / (dk["avg_ms"] * 1e-3)
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            tj = json.load(f)
        traffic = tj.get("kernels", {}).get(dominant, {}).get("hbm_bytes_per_launch")

    checked = None
    zh = z.cpu().numpy().view(np.uint64)
    if not args.no_check:
        ok = True
        for q in (2305843009213693951, 2305843009213693921, 2305843009213693907):
            ok &= (mod_q(a_h, q) * mod_q(b_h, q)) % q == mod_q(zh, q)
        checked = bool(ok)
        if not ok:
            print("bench.py: PRODUCT MISMATCH", file=sys.stderr)

    B, F = product_model(W)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(W, a_h, b_h, zh)
        except Exception as e:  # oracle missing: report null
            print("# cpu baseline unavailable: %s" % e, file=sys.stderr)
    fops = KERNEL_FP64_OPS.get(dominant)
    fp64_achieved = fops * W["ref_np"] * W["ref_T"] / (dk["avg_ms"] * 1e-3) if fops else None

    value = an * bn / (ms * 1e-3)
    out = {
        "metric": "limb-products/s batched vs plain GPU vs CPU, bit-exact; % HBM roofline",
        "value": value,
        "unit": "limb-products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": W["desc"],
            "an": an, "bn": bn, "seed": W["seed"],
            "plan": plan,
            "bit_exact_mod_q_check": checked,
            "product_model_bytes": B,
            "product_model_hbm_frac": (B / (ms * 1e-3)) / HBM_PEAK,
            "product_model_fp64_lane_ops": F,
            "product_model_fp64_frac": (F / (ms * 1e-3)) / FP64_LANE_OPS_PEAK,
            "kernels_avg_ms": {k: round(v["avg_ms"], 5) for k, v in per_kernel.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "fp64_lane_ops_per_s": fp64_achieved,
            "fp64_peak_lane_ops_per_s": FP64_LANE_OPS_PEAK,
            "fp64_frac": fp64_achieved / FP64_LANE_OPS_PEAK if fp64_achieved else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
