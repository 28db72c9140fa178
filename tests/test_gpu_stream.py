"""GPU parity of the full mixed stream's device side: operands written on the
device equal the oracle's restatement of the generator, the device
evaluation equals the oracle's checkers and Python integers, and a chunk of
the stream (its first requests, run through the graph-captured batch) passes
the A(x) B(x) = Z(x) mod m check, with its small products also compared
whole against the oracle.  A product corrupted on purpose (one limb changed
in HBM after the batch ran) is reported by the check."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import to_int

pytestmark = pytest.mark.gpu


def _fill(torch, kind, key, count, n=0, w=1, bits=0):
    import flint_amd as fa
    words = count * (w if kind == 2 else 1)
    t = torch.zeros(max(words, 1), dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    F = (fa.FillSpec * 1)()
    F[0].dst, F[0].count, F[0].key, F[0].kind = t.data_ptr(), count, key, kind
    F[0].n, F[0].w, F[0].bits = n, w, bits
    fa.stream_fill_device(F)
    return t, t.cpu().numpy().view(np.uint64)[:words]


@pytest.mark.parametrize("count", [1, 5, 1023, 1024, 1025, 70001])
def test_device_fill_matches_oracle_integer_and_nmod(count):
    import torch
    key = 0x1234567 + count
    _, got = _fill(torch, 0, key, count)
    assert np.array_equal(got, O.stream_operand(0, key, count))
    for n in (3, 1000003, (1 << 61) - 1, 2**63 - 25):
        _, got = _fill(torch, 1, key ^ n, count, n=n)
        assert np.array_equal(got, O.stream_operand(1, key ^ n, count, n=n)), n
        assert int(got.max()) < n


@pytest.mark.parametrize("bits", [1, 63, 64, 127, 128, 1000, 4096])
def test_device_fill_matches_oracle_fmpz(bits):
    import torch
    key = 0xABCDEF ^ bits
    w = bits // 64 + 1
    for ww in (w, w + 2):
        _, got = _fill(torch, 2, key, 3000, w=ww, bits=bits)
        want = O.stream_operand(2, key, 3000, w=ww, bits=bits)
        assert np.array_equal(got, want), (bits, ww)
    vals = []
    for i in range(300):
        v = to_int(want[i * ww:(i + 1) * ww])
        if v >> (64 * ww - 1):
            v -= 1 << (64 * ww)
        vals.append(v)
    assert all(-(1 << bits) < v < (1 << bits) for v in vals)
    assert min(vals) < 0 < max(vals) or bits == 1


def test_device_eval_matches_checkers():
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(11)
    cases = []
    for count in (1, 17, 4096, 4097, 100000):
        a = rng.integers(0, 2**63, count, dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, count, dtype=np.uint64)
        for m in (2, 3, 65537, (1 << 61) - 1, 2**63 - 25, 2**64 - 59):
            cases.append((a, 1, 0, (1 << 64) % m, m, O.mpn_mod_q(a, m)))
            x = int(rng.integers(0, 2**62))
            cases.append((a, 1, 0, x, m, O.nmod_poly_eval(a, x, m)))
    for w in (1, 3):
        c = rng.integers(0, 2**63, 999 * w, dtype=np.uint64) * np.uint64(2)
        c[w - 1::w] |= np.uint64(1 << 63)      # every coefficient negative
        c[5 * w + w - 1] = np.uint64(7)         # one positive
        for m in ((1 << 61) - 1, 1000003):
            x = 123456789 % m
            want = O.fmpz_poly_eval_mod_q(c, w, x, m)
            # Python integers
            vals = []
            for i in range(999):
                v = to_int(c[i * w:(i + 1) * w])
                vals.append(v - (1 << (64 * w)) if v >> (64 * w - 1) else v)
            assert want == sum(v * pow(x, i, m) for i, v in enumerate(vals)) % m
            cases.append((c, w, 1, x, m, want))
    E = (fa.EvalSpec * len(cases))()
    keep = []
    for j, (arr, w, sg, x, m, _) in enumerate(cases):
        t = torch.from_numpy(arr.view(np.int64)).to("cuda:0")
        keep.append(t)
        E[j].src, E[j].count, E[j].w, E[j].x, E[j].m, E[j].is_signed = \
            t.data_ptr(), len(arr) // w, w, x, m, sg
    torch.cuda.synchronize()
    got = fa.eval_device(E)
    for j, case in enumerate(cases):
        assert int(got[j]) == case[5], (j, len(case[0]), case[4])


def test_stream_chunk_checks_and_small_products():
    import torch
    import flint_amd as fa
    import flint_amd.workloads as W
    H = fa.stream_headers(240)
    # the headers of the oracle's restatement
    ref = O.stream_headers(240)
    for h, r in zip(H, ref):
        assert [h.kind, h.bits, h.an, h.bn, h.n, h.key[0], h.key[1]] == [int(v) for v in r]
    idx = [k for k in range(240) if W.stream_header_workspace(H[k]) < (1 << 30)]
    print("chunk of", len(idx), "requests", flush=True)
    ch = W.StreamChunk(H, idx, torch, "cuda:0")
    s = torch.cuda.Stream()
    import ctypes
    ch.run(ctypes.c_void_p(s.cuda_stream))
    s.synchronize()
    assert ch.check() == []
    # small products whole against the oracle, operands from the restatement
    done = 0
    for j, k in enumerate(idx):
        h = H[k]
        if done >= 12 or h.an * h.bn > 4e6:
            continue
        a_w, b_w, z_w, w, zw = W.stream_words(h)
        base = (ch.ptrs[j][2] - ch.arena.data_ptr()) // 8
        got = ch.arena[base:base + z_w].cpu().numpy().view(np.uint64)
        a = O.stream_operand(h.kind, h.key[0], h.an, n=h.n, w=w, bits=h.bits)
        b = O.stream_operand(h.kind, h.key[1], h.bn, n=h.n, w=w, bits=h.bits)
        if h.kind == 0:
            want = O.mpn_mul(a, b)
        elif h.kind == 1:
            want = O.nmod_poly_mul_mid(a, b, 0, h.an + h.bn - 1, h.n)
        else:
            ok, want = O.fmpz_poly_mul_mid(a, w, b, w, 0, h.an + h.bn - 1, zw)
            if not ok:
                continue
        assert np.array_equal(got, want), k
        done += 1
    assert done >= 6
    # one limb changed in HBM is caught, and only that request
    j = next(j for j, k in enumerate(idx) if H[k].kind == 0 and H[k].an + H[k].bn > 100)
    base = (ch.ptrs[j][2] - ch.arena.data_ptr()) // 8
    torch.cuda.synchronize()
    ch.arena[base + 50] ^= 1 << 20
    torch.cuda.synchronize()
    assert ch.check() == [idx[j]]
    ch.close()
