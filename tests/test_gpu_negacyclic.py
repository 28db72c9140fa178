"""GPU parity of products mod 2^N + 1 (sd_fft_mpn_mulmod_2expp1,
negacyclic.c; the pointwise stage of FLINT's Schoenhage-Strassen
convolutions): random normalized operands and the edge values 0, 1, 2^N - 1
and 2^N (= -1), whose products hit both branches of the fold (D = lo - hi
negative or not, the results 2^N and 1), against Python integers.  The
device entry point on one pointer for both operands takes the squaring
route."""
import numpy as np
import pytest

from tests.helpers import from_int, to_int

pytestmark = pytest.mark.gpu


def _mod_fermat(x, N):
    """x mod 2^N + 1 for 0 <= x <= 2^(2N) by 2^N = -1 (no long division)"""
    r = (x & ((1 << N) - 1)) - (x >> N)
    return r + (1 << N) + 1 if r < 0 else r


def _norm_rand(rng, n):
    N = 64 * n
    if rng.integers(0, 8) == 0:
        return 1 << N
    return int.from_bytes(rng.bytes(8 * n), "little")


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1000, 33000, 100000])
def test_mulmod_2expp1_matches_integers(n):
    import flint_amd as fa
    rng = np.random.default_rng(n)
    N = 64 * n
    edges = [0, 1, (1 << N) - 1, 1 << N, 2]
    pairs = [(u, v) for u in edges for v in edges]
    pairs += [(_norm_rand(rng, n), _norm_rand(rng, n)) for _ in range(6)]
    if n > 1000:
        pairs = pairs[::3]
    for u, v in pairs:
        x, y = from_int(u, n + 1), from_int(v, n + 1)
        z = fa.mpn_mulmod_2expp1(x, y, n)
        assert to_int(z) == _mod_fermat(u * v, N), (n, u.bit_length(), v.bit_length())
    # squares (one operand array)
    u = _norm_rand(rng, n)
    x = from_int(u, n + 1)
    assert to_int(fa.mpn_mulmod_2expp1(x, x, n)) == _mod_fermat(u * u, N)


def test_mulmod_2expp1_device_square_and_rejects():
    import torch
    import flint_amd as fa
    n = 4096
    rng = np.random.default_rng(3)
    u = int.from_bytes(rng.bytes(8 * n), "little")
    x = from_int(u, n + 1)
    dx = torch.from_numpy(x.view(np.int64)).to("cuda:0")
    dz = torch.zeros(n + 1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()   # the fill runs on torch's stream, the call on s
    s = torch.cuda.Stream()
    assert fa.lib().fa_mpn_mulmod_2expp1_device(dz.data_ptr(), dx.data_ptr(), dx.data_ptr(), n,
                                                s.cuda_stream) == 0
    s.synchronize()
    assert to_int(dz.cpu().numpy().view(np.uint64)) == _mod_fermat(u * u, 64 * n)
    bad = x.copy()
    bad[n] = 2
    with pytest.raises(fa.FlintAmdError):
        fa.mpn_mulmod_2expp1(bad, x, n)


@pytest.mark.parametrize("n,m", [(1, 0), (5, 16), (1000, 0), (4096, 64), (33000, 0)])
def test_mulmod_2expp1_cached_transform(n, m):
    """the context form (sd_fft_mpn_mulmod_2expp1_ctx_init / _transform /
    _mul_cached, the precached Schoenhage-Strassen shape): y's transform
    cached once and reused for several x, edge values included, and the
    one-shot product with the context, against Python integers"""
    import torch
    import flint_amd as fa
    rng = np.random.default_rng(n)
    N = 64 * n
    C = fa.Mulmod2expp1Ctx(N, m)
    assert (C.N, C.n, C.m) == (N, n, m)
    size = C.transformed_size()
    assert size > 0
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    for v in (_norm_rand(rng, n), 1 << N, (1 << N) - 1):
        F = torch.empty(size, dtype=torch.float64, device=dev)
        dy = torch.from_numpy(from_int(v, n + 1).view(np.int64)).to(dev)
        torch.cuda.synchronize()
        C.transform(F.data_ptr(), dy.data_ptr(), s.cuda_stream)
        xs = [_norm_rand(rng, n) for _ in range(3)] + [0, 1, 1 << N]
        for u in xs:
            print(n, m, u.bit_length(), v.bit_length(), flush=True)
            dx = torch.from_numpy(from_int(u, n + 1).view(np.int64)).to(dev)
            dz = torch.zeros(n + 1, dtype=torch.int64, device=dev)
            torch.cuda.synchronize()
            C.mul_cached(dz.data_ptr(), dx.data_ptr(), F.data_ptr(), s.cuda_stream)
            s.synchronize()
            assert to_int(dz.cpu().numpy().view(np.uint64)) == _mod_fermat(u * v, N)
        dz = torch.zeros(n + 1, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        C.mul(dz.data_ptr(), dx.data_ptr(), dy.data_ptr(), s.cuda_stream)
        s.synchronize()
        assert to_int(dz.cpu().numpy().view(np.uint64)) == _mod_fermat(xs[-1] * v, N)
    for badN, badm in ((100, 0), (64 * 5, 8), (64 * 5, 48), (64, 128)):
        with pytest.raises(fa.FlintAmdError):
            fa.Mulmod2expp1Ctx(badN, badm)
