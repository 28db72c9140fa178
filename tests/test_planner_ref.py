"""CPU checks of the planner's FP64 reference network and of the plan space.

The reference backward pass (oracle/mlp_ref.py) is pinned by central finite
differences; the candidate plans are pinned by the plan constraints they
must satisfy (exactness M >= min(alen, blen) 2^(2 bits), chunk sizes in
steps of 8, the automatic plan among them)."""
import numpy as np

from oracle import mlp_ref as R


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.3, R.NPARAM)
    X = rng.normal(0, 1, (7, 16))
    T = rng.normal(0, 1, (7, 64))
    M = (rng.random((7, 64)) < 0.7).astype(float)
    _, g = R.loss_grad(p, X, T, M)
    idx = rng.choice(R.NPARAM, 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros(R.NPARAM)
        e[k] = h
        fd = (R.loss_grad(p + e, X, T, M)[0] - R.loss_grad(p - e, X, T, M)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_candidates_satisfy_plan_constraints():
    import flint_amd as fa
    from oracle import oracle as O
    primes = O.default_primes()
    for an, bn in [(16, 16), (1000, 900), (65536, 65536), (1 << 22, 1 << 21), (100000, 20)]:
        c = fa.mpn_candidates(an, bn)
        auto = fa.mpn_plan(an, bn)
        assert (auto["np"], auto["bits"]) in {(int(x), int(y)) for x, y, _ in c}
        for j, (np_, bits, tr) in enumerate(c):
            if not np_:
                continue
            assert j < 56 and np_ == j // 8 + 2 and tr == j % 2
            assert bits % 8 == 0 and 8 <= bits <= 200
            # the untruncated twin of a truncated candidate is valid too
            if tr:
                assert tuple(c[j - 1]) == (np_, bits, 0)
                zl = int(-(-64 * an // int(bits)) + -(-64 * bn // int(bits)) - 1)
                depth = max(4, (zl - 1).bit_length())
                assert 13 <= depth <= 23 and -(-zl // 4096) < 1 << (depth - 12)
            M = 1
            for p in primes[:np_]:
                M *= int(p)
            alen = -(-64 * an // bits)
            blen = -(-64 * bn // bits)
            assert M >= min(alen, blen) << (2 * bits)
