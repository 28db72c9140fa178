"""Concurrent use from several host threads (the reference's default context
is thread-local, default_ctx.c:14-15, so its callers may multiply from many
threads at once): each thread has its own stream, scratch and prime-group
streams here.  Four threads run a mix of products at the same time; every
result must equal the same product computed alone."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(fa, k):
    rng = np.random.default_rng(1000 + k)
    out = []
    a, b = fa.synth_mpn_pair(0x7100 + k, (1 << 22) + 17 * k, (1 << 21) + 5)
    out.append(("mpn", a, b))
    a, b = fa.synth_mpn_pair(0x7200 + k, 1 << 24, (1 << 23) + k)    # depth 24: prime groups
    out.append(("mpn", a, b))
    n = (1 << 62) - 57
    out.append(("nmod", rng.integers(0, n, 100000, dtype=np.uint64),
                rng.integers(0, n, 70000 + k, dtype=np.uint64), n))
    w = 2
    a = rng.integers(0, 2**64, 3000 * w, dtype=np.uint64, endpoint=False)
    b = rng.integers(0, 2**64, 2500 * w, dtype=np.uint64, endpoint=False)
    a[1::2] &= np.uint64((1 << 55) - 1)   # 119-bit signed coefficients
    b[1::2] |= np.uint64(0xff80000000000000)
    out.append(("fmpz", a, b, w))
    return out


def _run(fa, job):
    if job[0] == "mpn":
        return fa.mpn_mul(job[1], job[2])
    if job[0] == "nmod":
        a, b, n = job[1], job[2], job[3]
        return fa.nmod_poly_mul_mid(a, b, 0, len(a) + len(b) - 1, n)
    a, b, w = job[1], job[2], job[3]
    zn = len(a) // w + len(b) // w - 1
    return fa.fmpz_poly_mul_mid(a, w, b, w, 0, zn, 5)


def test_concurrent_threads_match_sequential():
    import flint_amd as fa
    jobs = [_jobs(fa, k) for k in range(4)]
    want = [[_run(fa, j) for j in js] for js in jobs]
    got = [[None] * len(js) for js in jobs]
    errors = []

    def worker(k):
        try:
            for rep in range(2):
                for i, j in enumerate(jobs[k]):
                    z = _run(fa, j)
                    if rep == 1:
                        got[k][i] = z
        except Exception as e:   # reported below
            errors.append((k, repr(e)))

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not errors, errors
    for k in range(4):
        for i in range(len(jobs[k])):
            assert got[k][i] is not None and np.array_equal(got[k][i], want[k][i]), (k, i)
