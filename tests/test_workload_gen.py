"""CPU tests of the batched workloads' request generator and of the
full-size checkers the GPU parity tests use.

- The library's generator (fa_rng_*, used by flint_amd.workloads and the
  bench) and the oracle's independent restatement produce the same requests,
  draw for draw, for each workload.
- The draws follow their distributions' definitions: sizes in range,
  moduli prime with the drawn bit size, coefficients reduced / inside
  (-2^b, 2^b), top limbs nonzero.
- The checkers (or_mpn_mod_q, or_nmod_poly_eval, or_fmpz_poly_eval_mod_q)
  equal Python-integer evaluation on small inputs, including negative
  coefficients and values at the edges of their word sizes.
"""
import numpy as np
import pytest

import flint_amd.workloads as W
from oracle import oracle as O
from tests.helpers import to_int


def _signed_values(flat, w):
    out = []
    for i in range(len(flat) // w):
        v = to_int(flat[i * w:(i + 1) * w])
        if v >> (64 * w - 1):
            v -= 1 << (64 * w)
        out.append(v)
    return out


@pytest.mark.parametrize("workload,count", [("nmod4096", 12), ("fmpz256", 3), ("stream", 12)])
def test_library_and_oracle_generators_agree(workload, count):
    lib = W.make(workload, count)
    ref = O.workload_requests(workload, count)
    assert len(lib) == len(ref) == count
    for q, r in zip(lib, ref):
        assert q["kind"] == r[0]
        assert np.array_equal(q["a"], r[1]) and np.array_equal(q["b"], r[2])
        if r[0] == "nmod":
            assert q["n"] == r[3]
        if r[0] == "fmpz":
            assert q["aw"] == r[3] and q["bits"] == r[4]


def test_draws_follow_their_distributions():
    for q in W.make("nmod4096", 40):
        an, bn, n = len(q["a"]), len(q["b"]), q["n"]
        assert 64 <= an < 65536 and 64 <= bn < 65536
        assert 20 <= n.bit_length() <= 63 and O.lib().or_is_prime(n)
        assert int(q["a"].max()) < n and int(q["b"].max()) < n
    kinds = [q["kind"] for q in W.make("stream", 60)]
    assert {"mpn", "nmod", "fmpz"} == set(kinds)
    for q in W.make("fmpz256", 2):
        bits, w = q["bits"], q["aw"]
        assert 64 <= bits <= 4096 and w == bits // 64 + 1
        for v in _signed_values(q["a"], w)[:300] + _signed_values(q["b"], w)[:300]:
            assert -(1 << bits) < v < (1 << bits)
        # both signs occur and the magnitudes reach the top bit size
        va = _signed_values(q["a"], w)
        assert min(va) < 0 < max(va)
        assert max(abs(v) for v in va).bit_length() >= bits - 8
        assert q["abits"] == -max(abs(v) for v in va).bit_length()
    for q in W.make("stream", 30):
        if q["kind"] == "mpn":
            assert q["a"][-1] != 0 and q["b"][-1] != 0
            assert 16 <= len(q["a"]) < (1 << 22)


def test_signed_fill_edges():
    """bit sizes at limb boundaries: values stay in (-2^b, 2^b) and the
    extreme values -(2^b - 1) and 2^b - 1 are representable"""
    for bits in (63, 64, 65, 127, 128, 129):
        w = bits // 64 + 1
        r = W.Rng(bits)
        ro = O.Rng(bits)
        a = r.signed(2000, bits, w)
        assert np.array_equal(a, ro.signed(2000, bits, w))
        vals = _signed_values(a, w)
        assert all(-(1 << bits) < v < (1 << bits) for v in vals)
        assert min(vals) < -(1 << (bits - 2)) and max(vals) > (1 << (bits - 2))


def test_checkers_match_python_integers():
    rng = np.random.default_rng(4)
    qs = [(1 << 61) - 1, 1152921504606846883, 3, 2, 2**63 - 25, 2**64 - 59]
    for q in qs:
        for n in (1, 2, 7, 50):
            a = rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False)
            a[-1] = np.uint64(2**64 - 1)
            assert O.mpn_mod_q(a, q) == to_int(a) % q
        c = rng.integers(0, 2**64, 33, dtype=np.uint64, endpoint=False)
        x = int(rng.integers(0, 2**62))
        assert O.nmod_poly_eval(c, x, q) == sum(int(v) * pow(x, i, q) for i, v in enumerate(c)) % q
        for w in (1, 2, 3):
            vals = [int(v) for v in rng.integers(-2**62, 2**62, 17)]
            vals += [-(1 << (64 * w - 1)), (1 << (64 * w - 1)) - 1, 0, -1]
            flat = np.zeros(len(vals) * w, dtype=np.uint64)
            for i, v in enumerate(vals):
                flat[i * w:(i + 1) * w] = np.frombuffer(
                    (v % (1 << (64 * w))).to_bytes(8 * w, "little"), dtype=np.uint64)
            want = sum(v * pow(x, i, q) for i, v in enumerate(vals)) % q
            assert O.fmpz_poly_eval_mod_q(flat, w, x, q) == want


# ----------------------------------------------------------------------
# the full stream: headers and operand contents keyed per operand
# ----------------------------------------------------------------------

def test_stream_word_is_splitmix64():
    """SplitMix64's published outputs for seed 1234567, and the library's
    word function against the oracle's at large indices"""
    import flint_amd as fa
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [O.stream_word(1234567, j) for j in range(5)] == want
    assert [fa.stream_word(1234567, j) for j in range(5)] == want
    for key, j in ((0, 0), (2**64 - 1, 10**12), (0xF1A70005, 2**40 + 3)):
        assert fa.stream_word(key, j) == O.stream_word(key, j)


def test_stream_headers_agree_and_follow_their_distributions():
    import flint_amd as fa
    H = fa.stream_headers(3000)
    R = O.stream_headers(3000)
    kinds = [0, 0, 0]
    for h, r in zip(H, R):
        assert [h.kind, h.bits, h.an, h.bn, h.n, h.key[0], h.key[1]] == [int(v) for v in r]
        kinds[h.kind] += 1
        if h.kind == 0:
            assert 16 <= h.an < (1 << 22) and 16 <= h.bn < (1 << 22)
        elif h.kind == 1:
            assert 64 <= h.an < 65536 and 20 <= h.bits <= 63
            assert h.n.bit_length() == h.bits and O.lib().or_is_prime(h.n)
        else:
            assert 1024 <= h.an <= 16384 and 64 <= h.bits <= 4096
    assert abs(kinds[0] / 3000 - 0.5) < 0.04 and abs(kinds[1] / 3000 - 0.3) < 0.04


def test_stream_operands_from_their_words():
    """each kind's element mapping, recomputed from the words with Python
    integers"""
    key = 0x5EED
    a = O.stream_operand(0, key, 300)
    assert [int(v) for v in a] == [O.stream_word(key, j) for j in range(300)]
    n = 1000003
    c = O.stream_operand(1, key, 300, n=n)
    assert [int(v) for v in c] == [O.stream_word(key, j) * n >> 64 for j in range(300)]
    for bits in (1, 64, 100, 191):
        w = bits // 64 + 2
        tw = (bits + 64) // 64
        f = O.stream_operand(2, key, 50, w=w, bits=bits)
        for j, v in enumerate(_signed_values(f, w)):
            t = sum(O.stream_word(key, j * tw + k) << (64 * k) for k in range(tw))
            t &= (1 << (bits + 1)) - 1
            if t == (1 << (bits + 1)) - 1:
                t = 0
            assert v == t - ((1 << bits) - 1), (bits, j)
