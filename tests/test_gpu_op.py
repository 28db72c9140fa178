"""GPU parity of plans and transformed operands (fft_small.h:530-700,
op.c, the reference's t-op_mpn.c / t-op_nmod.c pattern): sums of products
accumulated in transform space, then one inverse transform and one export,
checked against Python integers and the oracle.

- Integers: sum_i a_i b_i with len_bound = K, including a square (op_sqr),
  operand sums (op_add before the product), submul and neg/add in the
  product domain, at transform depths on both sides of 12 (one LDS array, or
  column and row passes).
- nmod: a dot product of K polynomial pairs mod n, exported whole and as a
  window, against K oracle products summed mod n.
- The domain rules: a product of a product, addmul into a primal op and
  add of mixed domains are refused with errors.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import from_int, rand_limbs, to_int

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda:0")
    # the upload runs on torch's stream, the library's kernels on their own:
    # a pageable copy may still be in flight when .to() returns
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("an,bn,K", [(40, 30, 3), (700, 500, 4), (20000, 9000, 5),
                                     (150000, 150000, 3)])
def test_mpn_dot_product_in_transform_space(an, bn, K):
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(an + K)
    P = fa.Plan.mpn(an, bn, len_bound=4 * K)
    print("plan np=%d bits=%d depth=%d" % (P.s.np, P.s.bits, P.s.depth), flush=True)
    A = [rand_limbs(rng, an) for _ in range(K)]
    B = [rand_limbs(rng, bn) for _ in range(K)]
    dA = [_dev(torch, a) for a in A]
    dB = [_dev(torch, b) for b in B]
    opA = [fa.Op(P) for _ in range(K)]
    opB = [fa.Op(P) for _ in range(K)]
    for k in range(K):
        opA[k].fft_mpn(dA[k].data_ptr(), an)
        opB[k].fft_mpn(dB[k].data_ptr(), bn)
    Z = fa.Op(P)
    Z.mul(opA[0], opB[0])
    for k in range(1, K):
        Z.addmul(opA[k], opB[k])
    assert Z.domain == fa.OP_PRODUCT
    Z.ifft()
    zn = an + bn + 1
    dz = torch.zeros(zn, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()   # the fill runs on torch's stream, the library on its own
    Z.export_mpn(dz.data_ptr(), zn)
    torch.cuda.synchronize()
    want = sum(to_int(a) * to_int(b) for a, b in zip(A, B))
    assert to_int(dz.cpu().numpy().view(np.uint64)) == want % (1 << (64 * zn))

    # (a0 + a1) b0 - a1 b0 = a0 b0 (op_add before the product, submul)
    S = fa.Op(P)
    S.add(opA[0], opA[1])
    W = fa.Op(P)
    W.mul(S, opB[0])
    W.submul(opA[1], opB[0])
    W.ifft()
    W.export_mpn(dz.data_ptr(), zn)
    torch.cuda.synchronize()
    assert to_int(dz.cpu().numpy().view(np.uint64)) == to_int(A[0]) * to_int(B[0])

    # a0^2 via op_sqr, and a0 b0 + a1 b1 - a1 b1 through neg and add in the
    # product domain; a short export keeps the low limbs
    if an == bn:
        Q = fa.Op(P)
        Q.sqr(opA[0])
        Q.ifft()
        dq = torch.zeros(2 * an, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        Q.export_mpn(dq.data_ptr(), 2 * an)
        torch.cuda.synchronize()
        assert to_int(dq.cpu().numpy().view(np.uint64)) == to_int(A[0]) ** 2
    X = fa.Op(P)
    X.mul(opA[0], opB[0])
    X.addmul(opA[1], opB[1])
    Y = fa.Op(P)
    Y.mul(opA[1], opB[1])
    Y.neg(Y)
    X.add(X, Y)
    X.ifft()
    short = (an + bn) // 2
    ds = torch.zeros(short, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    X.export_mpn(ds.data_ptr(), short)
    torch.cuda.synchronize()
    assert to_int(ds.cpu().numpy().view(np.uint64)) == \
        (to_int(A[0]) * to_int(B[0])) % (1 << (64 * short))


@pytest.mark.parametrize("n,ln,K", [((1 << 61) - 1, 300, 4), (0x0003F00000000001, 5000, 3),
                                    (1000003, 40000, 6), (2**64 - 59, 2000, 2)])
def test_nmod_dot_product_in_transform_space(n, ln, K):
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(ln)

    def draw():
        if n < 2**63:
            return rng.integers(0, n, ln, dtype=np.uint64)
        return np.array([int.from_bytes(rng.bytes(8), "little") % n for _ in range(ln)],
                        dtype=np.uint64)
    A = [draw() for _ in range(K)]
    B = [draw() for _ in range(K)]
    zn = 2 * ln - 1
    P = fa.Plan.nmod(0, zn, zn, ln, K * ln, 2 * int(n).bit_length(), n)
    print("plan np=%d offset=%d depth=%d" % (P.s.np, P.s.offset, P.s.depth), flush=True)
    Z = fa.Op(P)
    ta, tb = fa.Op(P), fa.Op(P)
    for k in range(K):
        da, db = _dev(torch, A[k]), _dev(torch, B[k])
        ta.fft_nmod(da.data_ptr(), ln, n)
        tb.fft_nmod(db.data_ptr(), ln, n)
        if k == 0:
            Z.mul(ta, tb)
        else:
            Z.addmul(ta, tb)
        torch.cuda.synchronize()
    Z.ifft()
    dz = torch.zeros(zn, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    Z.export_nmod(dz.data_ptr(), n)
    lo, hi = zn // 3, zn // 3 + 101
    dw = torch.zeros(hi - lo, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    Z.export_nmod(dw.data_ptr(), n, lo, hi)
    torch.cuda.synchronize()
    want = np.zeros(zn, dtype=object)
    for k in range(K):
        want += O.nmod_poly_mul_mid(A[k], B[k], 0, zn, n).astype(object)
    want = np.array([int(v) % n for v in want], dtype=np.uint64)
    got = dz.cpu().numpy().view(np.uint64)
    assert np.array_equal(got, want)
    assert np.array_equal(dw.cpu().numpy().view(np.uint64), want[lo:hi])


def test_domain_rules_are_enforced():
    import torch
    import flint_amd as fa
    P = fa.Plan.mpn(100, 100, 2)
    a = _dev(torch, np.arange(1, 101, dtype=np.uint64))
    A, B, Z = fa.Op(P), fa.Op(P), fa.Op(P)
    A.fft_mpn(a.data_ptr(), 100)
    B.fft_mpn(a.data_ptr(), 100)
    Z.mul(A, B)
    with pytest.raises(fa.FlintAmdError):
        Z.mul(Z, A)                    # a product of a product
    with pytest.raises(fa.FlintAmdError):
        A.addmul(A, B)                 # addmul into a primal op
    with pytest.raises(fa.FlintAmdError):
        Z.add(Z, A)                    # mixed domains
    with pytest.raises(fa.FlintAmdError):
        A.ifft()                       # inverse of a primal op
    other = fa.Plan.mpn(5000, 5000, 2)
    C = fa.Op(other)
    with pytest.raises(fa.FlintAmdError):
        C.mul(A, B)                    # ops of another plan
    torch.cuda.synchronize()


@pytest.mark.parametrize("an,bn", [(30, 20), (3000, 2500), (70000, 40000)])
def test_signed_export(an, bn):
    """a0 b0 - a1 b1 and its negation through submul and the signed export
    (export_signed.c): |value| and the sign, against Python integers; a zero
    difference has sign 0"""
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(an)
    P = fa.Plan.mpn(an, bn, len_bound=2, signed=True)
    assert P.s.pad == 1
    A = [rand_limbs(rng, an) for _ in range(2)]
    B = [rand_limbs(rng, bn) for _ in range(2)]
    ops = []
    # the device operands stay referenced until the transforms (enqueued on
    # the library's stream) have read them
    dA = [_dev(torch, A[k]) for k in range(2)]
    dB = [_dev(torch, B[k]) for k in range(2)]
    for k in range(2):
        oa, ob = fa.Op(P), fa.Op(P)
        oa.fft_mpn(dA[k].data_ptr(), an)
        ob.fft_mpn(dB[k].data_ptr(), bn)
        ops.append((oa, ob))
    nl = -(-50 * P.s.np // 64)
    zn = (P.s.zn * P.s.bits + 64 * nl + 2 + 63) // 64 + 1
    dz = torch.zeros(zn, dtype=torch.int64, device="cuda:0")
    ds = torch.full((1,), 7, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    for first, second in ((0, 1), (1, 0), (0, 0)):
        Z = fa.Op(P)
        Z.mul(ops[first][0], ops[first][1])
        Z.submul(ops[second][0], ops[second][1])
        Z.ifft()
        Z.export_mpn_signed(dz.data_ptr(), zn, ds.data_ptr(), P.s.zn)
        torch.cuda.synchronize()
        want = to_int(A[first]) * to_int(B[first]) - to_int(A[second]) * to_int(B[second])
        got = to_int(dz.cpu().numpy().view(np.uint64))
        sign = int(ds.cpu()[0])
        assert (-got if sign else got) == want, (first, second)
        assert sign == (1 if want < 0 else 0)
    # an unsigned plan refuses the signed export
    Pu = fa.Plan.mpn(an, bn, len_bound=2)
    U = fa.Op(Pu)
    U.fft_mpn(dA[0].data_ptr(), an)
    U.sqr(U)
    U.ifft()
    with pytest.raises(fa.FlintAmdError):
        U.export_mpn_signed(dz.data_ptr(), zn, ds.data_ptr(), Pu.s.zn)


def _slot_values(a, b, bits):
    """the chunk convolution c_j = sum_{i+k=j} A_i B_k (A_i = bits
    [i bits, (i+1) bits) of a, bits a multiple of 8) by Kronecker
    substitution on Python integers"""
    cb = bits // 8
    alen, blen = -(-64 * len(a) // bits), -(-64 * len(b) // bits)
    S = (2 * bits + min(alen, blen).bit_length() + 8) // 8      # bytes per slot

    def spread(x, n):
        raw = np.frombuffer(np.asarray(x, dtype=np.uint64).tobytes(), dtype=np.uint8)
        raw = np.concatenate([raw, np.zeros(n * cb - len(raw), dtype=np.uint8)])
        out = np.zeros((n, S), dtype=np.uint8)
        out[:, :cb] = raw.reshape(n, cb)
        return int.from_bytes(out.tobytes(), "little")

    prod = spread(a, alen) * spread(b, blen)
    zlen = alen + blen - 1
    raw = prod.to_bytes(zlen * S + S, "little")
    return [int.from_bytes(raw[j * S:(j + 1) * S], "little") for j in range(zlen)]


def _slot_sum(c, bits, j0):
    """sum_{j < j0} c_j 2^(j bits) for 0 <= c_j < 2^(3 bits), as three
    concatenations of bits-wide fields"""
    cb, mask = bits // 8, (1 << bits) - 1
    total = 0
    for t in range(3):
        field = b"".join(((v >> (t * bits)) & mask).to_bytes(cb, "little") for v in c[:j0])
        total += int.from_bytes(field, "little") << (t * bits)
    return total


def _signed_trunc_expected(A, B, bits, np_, nslots, lo, zn, slots=None):
    """export_signed.c:367-606 on Python integers: V' = the sum of the slots
    j >= j0 of a0 b0 - a1 b1 (j0 = the first slot that can reach limb lo,
    rounded down to a multiple of 256 on plans of more than two primes);
    z = limbs [lo, lo + zn) of |V'|, rounded up for negative V' on those
    plans (the window of ~W + 1), down on two-prime plans"""
    from tests.test_oracle_window import _coeff_len
    n = _coeff_len(np_)
    j0 = 0
    if 64 * lo > 64 * n + 1:
        j0 = -(-(64 * lo - 64 * n - 1) // bits)
    j0 = min(j0, nslots)
    if np_ > 2:
        j0 = j0 // 256 * 256
    if slots is None:
        slots = [_slot_values(A[k], B[k], bits) for k in range(2)]
    V = to_int(A[0]) * to_int(B[0]) - to_int(A[1]) * to_int(B[1])
    V -= _slot_sum(slots[0], bits, j0) - _slot_sum(slots[1], bits, j0)
    T = 1 << (64 * lo)
    mag = abs(V)
    if V < 0 and np_ > 2:
        z = -(-mag // T)
    else:
        z = mag // T
    return z % (1 << (64 * zn)), int(V < 0), V


@pytest.mark.parametrize("np_", [2, 0])
@pytest.mark.parametrize("an,bn", [(30, 20), (3000, 2500), (9000, 7000)])
def test_signed_export_trunc(an, bn, np_):
    """fft_small_export_mpn_signed_trunc: limbs [lo, lo + zn) of the signed
    sum a0 b0 - a1 b1 with the reference's slot cut and rounding, on a
    two-prime plan and an automatic (wider) one, both signs, against the
    rule restated on Python integers; lo = 0 is the exact export"""
    import torch
    import flint_amd as fa
    from tests.test_oracle_window import _coeff_len
    rng = np.random.default_rng(an + np_)
    P = fa.Plan.mpn(an, bn, len_bound=2, signed=True, np=np_)
    assert (P.s.np == 2) == (np_ == 2)
    A = [rand_limbs(rng, an) for _ in range(2)]
    B = [rand_limbs(rng, bn) for _ in range(2)]
    dA = [_dev(torch, A[k]) for k in range(2)]
    dB = [_dev(torch, B[k]) for k in range(2)]
    ops = []
    for k in range(2):
        oa, ob = fa.Op(P), fa.Op(P)
        oa.fft_mpn(dA[k].data_ptr(), an)
        ob.fft_mpn(dB[k].data_ptr(), bn)
        ops.append((oa, ob))
    nslots, bits = P.s.zn, P.s.bits
    need = -(-(nslots * bits + 64 * _coeff_len(P.s.np) + 2) // 64)
    ds = torch.full((1,), 7, dtype=torch.int32, device="cuda:0")
    for first, second in ((0, 1), (1, 0)):
        Z = fa.Op(P)
        Z.mul(ops[first][0], ops[first][1])
        Z.submul(ops[second][0], ops[second][1])
        Z.ifft()
        Ao, Bo = [A[first], A[second]], [B[first], B[second]]
        slots = [_slot_values(Ao[k], Bo[k], bits) for k in range(2)]
        for lo in sorted({0, 1, 3, need // 3, need // 2, need - 6, need - 2}):
            zn = need - lo + 1
            print(an, bn, np_, first, lo, zn, flush=True)
            dz = torch.zeros(zn, dtype=torch.int64, device="cuda:0")
            torch.cuda.synchronize()
            Z.export_mpn_signed_trunc(dz.data_ptr(), zn, ds.data_ptr(), nslots, lo)
            torch.cuda.synchronize()
            got = to_int(dz.cpu().numpy().view(np.uint64))
            sign = int(ds.cpu()[0])
            want, wsign, V = _signed_trunc_expected(Ao, Bo, bits, P.s.np, nslots, lo, zn, slots)
            assert (got, sign) == (want, wsign), (first, lo)
            # within two units of limb lo of the exact signed value
            exact = abs(V) >> (64 * lo)
            assert abs(got - exact) <= 2
            if lo == 0:
                assert (-got if sign else got) == V
    # too few limbs for the value is refused
    with pytest.raises(fa.FlintAmdError):
        Z.export_mpn_signed_trunc(dz.data_ptr(), 1, ds.data_ptr(), nslots, 0)


@pytest.mark.parametrize("nn_req,K", [(100, 1), (1500, 3), (123, 2), (7999, 1), (50000, 2)])
def test_mpn_cyclic_plan(nn_req, K):
    """fft_small_plan_init_mpn_cyclic as t-mpn_cyclic.c checks it: the
    rounded-up length tiles exactly (bits 2^depth = 64 nn), and a sum of K
    products of operands of up to nn limbs, exported and folded, is the sum
    mod B^nn - 1; squares, an operand of all ones (B^nn - 1 = 0) and short
    operands included"""
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(nn_req)
    P, nn = fa.Plan.mpn_cyclic(nn_req, K)
    bits, depth = P.s.bits, P.s.depth
    assert nn >= nn_req and bits << depth == 64 * nn
    zn = nn + (2 * bits + depth + K.bit_length()) // 64 + 1
    mod = (1 << (64 * nn)) - 1
    for case in ("full", "short", "square", "ones"):
        A, B, want = [], [], 0
        for k in range(K):
            an = nn if case != "short" else int(rng.integers(1, nn + 1))
            bn = nn if case != "short" else int(rng.integers(1, nn + 1))
            a = rand_limbs(rng, an)
            if case == "ones" and k == 0:
                a[:] = np.uint64(2**64 - 1)
            b = a.copy() if case == "square" else rand_limbs(rng, bn)
            A.append(a)
            B.append(b)
            want += to_int(a) * to_int(b)
        print(nn_req, K, nn, case, flush=True)
        dA = [_dev(torch, x) for x in A]
        dB = [_dev(torch, x) for x in B]
        Z, T, oa, ob = fa.Op(P), fa.Op(P), fa.Op(P), fa.Op(P)
        for k in range(K):
            oa.fft_mpn(dA[k].data_ptr(), len(A[k]))
            if case == "square":
                (Z if k == 0 else T).sqr(oa)
                if k:
                    Z.add(Z, T)
            else:
                ob.fft_mpn(dB[k].data_ptr(), len(B[k]))
                if k == 0:
                    Z.mul(oa, ob)
                else:
                    Z.addmul(oa, ob)
            torch.cuda.synchronize()
        Z.ifft()
        dz = torch.zeros(zn, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        Z.export_mpn(dz.data_ptr(), zn)
        torch.cuda.synchronize()
        got = to_int(dz.cpu().numpy().view(np.uint64))
        assert got % mod == want % mod, case
    with pytest.raises(fa.FlintAmdError):
        fa.Plan.mpn_cyclic(nn_req, 2**64 - 1)       # no plan: refused


@pytest.mark.parametrize("depth,n", [(4, 65537), (9, (1 << 61) - 1), (13, 2**64 - 59),
                                     (16, 1000003)])
def test_nmod_cyclic_plan(depth, n):
    """fft_small_plan_init_nmod_cyclic: the product of two length-N
    polynomials mod x^N - 1 over Z/nZ, coefficients [0, zh), against the
    oracle's linear product folded"""
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(depth)
    N = 1 << depth

    def draw(m):
        return np.array([int.from_bytes(rng.bytes(8), "little") % n for _ in range(m)],
                        dtype=np.uint64)
    a, b = draw(N), draw(N - 3)
    for zh in (N, N // 2 + 1):
        P = fa.Plan.nmod_cyclic(depth, zh, N, 2 * int(n).bit_length(), n)
        print(depth, n, zh, P.s.np, flush=True)
        da, db = _dev(torch, a), _dev(torch, b)
        ta, tb, Z = fa.Op(P), fa.Op(P), fa.Op(P)
        ta.fft_nmod(da.data_ptr(), N, n)
        tb.fft_nmod(db.data_ptr(), N - 3, n)
        Z.mul(ta, tb)
        Z.ifft()
        dz = torch.zeros(zh, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        Z.export_nmod(dz.data_ptr(), n)
        torch.cuda.synchronize()
        full = O.nmod_poly_mul_mid(a, b, 0, 2 * N - 4, n)
        cyc = [(int(full[k]) + (int(full[k + N]) if k + N < len(full) else 0)) % n
               for k in range(N)]
        assert [int(x) for x in dz.cpu().numpy().view(np.uint64)] == cyc[:zh]
