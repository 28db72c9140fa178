"""GPU parity of the batched workloads at their full sizes, against checkers
that share nothing with the HIP path.

- nmod4096: all 4096 products of the seeded workload (lengths log-uniform in
  [2^6, 2^16], a random prime modulus of 20..63 bits per pair) through one
  graph-captured device batch; every output equals the oracle's product
  (t-nmod_poly_mul.c:244 compares against the classical product the same
  way).  A second replay of the graph gives the same words.
- fmpz256: all 256 products (lengths uniform in [1024, 16384], coefficient
  bits uniform in [64, 4096]; nearly all beyond eight primes, computed by
  Kronecker substitution inside the batch, as FLINT's _fmpz_poly_mul does)
  checked by evaluation: z(x) = a(x) b(x) mod q at three (61-bit q, x)
  pairs, the coefficients read as two's complement (the check of
  t-fmpz_poly_mul.c:114 at a size where exact products are too slow).
- stream: the first 2000 requests of the seeded mixed stream, integer plans
  from the learned planner, every product checked mod q (integers) or by
  evaluation (nmod, fmpz).
Each workload runs in batches of at most 32 GB of device workspace.
"""
import ctypes
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

QX = [((1 << 61) - 1, 0x1234567890ABCDE), (1152921504606846883, 987654321987654321),
      (4611686018427387847, 0x0FEDCBA987654321)]
THREADS = 16


def _run_batches(reqs, torch, ks=True):
    """device batches over `reqs`; returns (host outputs, statuses)"""
    import flint_amd.workloads as W
    dev = torch.device("cuda:0")
    outs, status = [None] * len(reqs), [None] * len(reqs)
    for ch in W.chunks(reqs, 32 << 30):
        B, touts, keep = W.device_batch([reqs[k] for k in ch], torch, dev, ks=ks)
        R = keep[-1]
        B.launch()
        torch.cuda.synchronize()
        first = [t.cpu().numpy().view(np.uint64).copy() for t in touts]
        for t in touts:
            t.zero_()
        B.launch()
        torch.cuda.synchronize()
        for j, k in enumerate(ch):
            again = touts[j].cpu().numpy().view(np.uint64)
            assert np.array_equal(first[j], again), ("replay differs", k)
            outs[k] = first[j]
            status[k] = R[j].status
        B.close()
        del touts, keep
        torch.cuda.empty_cache()
        print("batch of %d requests done" % len(ch), flush=True)
    return outs, status


def _check_nmod(q, z):
    zn = len(q["a"]) + len(q["b"]) - 1
    want = O.nmod_poly_mul_mid(q["a"], q["b"], 0, zn, q["n"])
    return np.array_equal(z[:zn], want)


def _eval_ok(q, z):
    """evaluation check of one product at the (q, x) pairs"""
    if q["kind"] == "mpn":
        for p, _ in QX:
            if O.mpn_mod_q(z, p) != O.mpn_mod_q(q["a"], p) * O.mpn_mod_q(q["b"], p) % p:
                return False
        return True
    if q["kind"] == "nmod":
        n = q["n"]
        zn = len(q["a"]) + len(q["b"]) - 1
        for _, x in QX:
            if O.nmod_poly_eval(z[:zn], x, n) != \
                    O.nmod_poly_eval(q["a"], x, n) * O.nmod_poly_eval(q["b"], x, n) % n:
                return False
        return True
    w, zw = q["aw"], q["zw"]
    zn = len(q["a"]) // w + len(q["b"]) // q["bw"] - 1
    for p, x in QX:
        lhs = O.fmpz_poly_eval_mod_q(z[: zn * zw], zw, x, p)
        rhs = O.fmpz_poly_eval_mod_q(q["a"], w, x, p) * \
            O.fmpz_poly_eval_mod_q(q["b"], q["bw"], x, p) % p
        if lhs != rhs:
            return False
    return True


def test_nmod4096_full_batch_vs_oracle():
    import torch
    import flint_amd.workloads as W
    reqs = W.make("nmod4096")
    assert len(reqs) == 4096
    print("max length", max(max(len(q["a"]), len(q["b"])) for q in reqs), flush=True)
    outs, status = _run_batches(reqs, torch)
    assert all(s == 1 for s in status)
    t0 = time.time()
    with ThreadPoolExecutor(THREADS) as ex:
        ok = list(ex.map(lambda k: _check_nmod(reqs[k], outs[k]), range(len(reqs))))
    print("oracle check %.1fs" % (time.time() - t0), flush=True)
    bad = [k for k, v in enumerate(ok) if not v]
    assert not bad, bad[:10]


def test_fmpz256_full_batch_by_evaluation():
    import torch
    import flint_amd.workloads as W
    reqs = W.make("fmpz256")
    assert len(reqs) == 256
    outs, status = _run_batches(reqs, torch, ks=True)
    assert all(s == 1 for s in status)
    # the evaluation check catches a wrong coefficient: flip one word of a
    # copy and see it fail
    z = outs[0].copy()
    z[len(z) // 2] ^= np.uint64(1 << 17)
    assert not _eval_ok(reqs[0], z)
    with ThreadPoolExecutor(THREADS) as ex:
        ok = list(ex.map(lambda k: _eval_ok(reqs[k], outs[k]), range(len(reqs))))
    bad = [k for k, v in enumerate(ok) if not v]
    assert not bad, bad[:10]


def test_stream_first_2000_learned_planner():
    import torch
    import flint_amd.workloads as W
    from flint_amd import planner as PL
    t0 = time.time()
    reqs = W.make("stream", 2000)
    print("generated in %.1fs" % (time.time() - t0), flush=True)
    net = PL.load_planner()
    reqs = W.with_plans(reqs, net)
    kinds = {k: sum(q["kind"] == k for q in reqs) for k in ("mpn", "nmod", "fmpz")}
    print(kinds, flush=True)
    outs, status = _run_batches(reqs, torch, ks=True)
    assert all(s == 1 for s in status)
    with ThreadPoolExecutor(THREADS) as ex:
        ok = list(ex.map(lambda k: _eval_ok(reqs[k], outs[k]), range(len(reqs))))
    bad = [k for k, v in enumerate(ok) if not v]
    assert not bad, [(k, reqs[k]["kind"]) for k in bad[:10]]
    # the small integer products also against the exact product
    small = [k for k, q in enumerate(reqs) if q["kind"] == "mpn" and
             len(q["a"]) + len(q["b"]) < 20000][:200]
    for k in small:
        assert np.array_equal(outs[k], O.mpn_mul(reqs[k]["a"], reqs[k]["b"])), k
