"""GPU test of fa_mpn_mul_device_noalloc, the integer product entry point
that allocates nothing (scratch in a static device array): two calls in a
row, the second smaller, give the words fa_mpn_mul_device gives (and the
oracle's product); a product whose scratch exceeds the array is refused."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_limbs

pytestmark = pytest.mark.gpu


def test_noalloc_matches_device_entry_point():
    import torch
    import flint_amd as fa
    L = fa.lib()
    assert L.fa_init() == 0
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    rng = np.random.default_rng(77)
    for an, bn in ((600000, 500000), (3000, 2000), (1, 1)):
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        da = torch.from_numpy(a.view(np.int64)).to(dev)
        db = torch.from_numpy(b.view(np.int64)).to(dev)
        z1 = torch.zeros(an + bn, dtype=torch.int64, device=dev)
        z2 = torch.full((an + bn,), -1, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        fa.mpn_mul_device(z1.data_ptr(), da.data_ptr(), an, db.data_ptr(), bn, s.cuda_stream)
        rc = L.fa_mpn_mul_device_noalloc(z2.data_ptr(), da.data_ptr(), an, db.data_ptr(), bn,
                                         s.cuda_stream)
        assert rc == 0, L.fa_last_error()
        s.synchronize()
        want = z1.cpu().numpy().view(np.uint64)
        got = z2.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, want), (an, bn)
        if an + bn < 20000:
            assert np.array_equal(got, O.mpn_mul(a, b))
    # above the static array: a clear error, no launch
    big = 4 << 20
    dz = torch.zeros(2, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    rc = L.fa_mpn_mul_device_noalloc(dz.data_ptr(), dz.data_ptr(), big, dz.data_ptr(), big,
                                     s.cuda_stream)
    assert rc == -4           # FA_E_RANGE
    assert b"scratch" in L.fa_last_error()
    assert L.fa_noalloc_scratch_bytes() == 64 << 20
