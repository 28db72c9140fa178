"""CPU checks of bench.py's contract and of the geometry its byte model uses.

- The reference-plan geometry bench.py prices its roofline on (np, bits,
  ztrunc, depth per workload) is the plan the oracle's restatement of
  mpn_ctx_best_profile picks for that shape on one thread.
- The 2^24 x 2^24 byte model is 8.116 GB (ten array passes), and the
  per-kernel bytes of the fused pipeline come to two passes less.
- bench.py refuses more than one GPU with exit status 2 and says why.
"""
import os
import subprocess
import sys

import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


@pytest.mark.parametrize("key", ["mpn_2p24", "mpn_65536"])
def test_priced_geometry_is_the_reference_plan(key):
    W = _bench().WORKLOADS[key]
    pl = O.mpn_plan(W["an"], W["bn"], 1)
    assert (pl["np"], pl["bits"], pl["ztrunc"], pl["depth"]) == (
        W["ref_np"], W["ref_bits"], W["ref_T"], W["ref_depth"])


def test_byte_model_of_the_2p24_product():
    b = _bench()
    W = b.WORKLOADS["mpn_2p24"]
    B, _ = b.product_model(W)
    # 10 array passes of 8 bytes over np * ztrunc points, limbs in and out
    assert B == 8 * 6 * 15790336 * 10 + 8 * (2 ** 25) * 2 == 8116232192
    per = {k: b.kernel_bytes(k, W) for k in ("fa_k_col_fwd16w", "fa_k_row_conv16",
                                            "fa_k_col_inv16w", "fa_k_crt_fp")}
    # the fused row pass (forward rows of a and b, pointwise, inverse rows)
    # leaves out two of the model's ten passes
    assert sum(per.values()) == B - 2 * 8 * 6 * 15790336 == 6600359936


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
