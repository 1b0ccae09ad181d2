"""Every pass split the autotuner may choose must give the same bits: each
candidate split (single-transform and batch tables) is run and compared on
the device with the output of the default split, itself checked against the
oracle."""
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host
from zk_gpu_amd.autotune import TILE_LOG, candidates

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [13, 14, 15, 16, 17, 18, 19, 20])
@pytest.mark.parametrize("kind", ["single", "batch"])
def test_all_candidate_splits(torch_cuda, k, kind):
    torch = torch_cuda
    n = 1 << k
    x = O.gen(600 + k, n)
    ref = {}
    for op in ("fwd", "coset_inv"):
        ref[op] = to_dev(torch, O.run_op(op, x, k, threads=8))
    d_in = to_dev(torch, x)
    d_out = dev_empty(torch, n)
    plan_kind = Z.PLAN_SINGLE if kind == "single" else Z.PLAN_BATCH
    bad = []
    try:
        for digits, cols in candidates(k, tile_log_max=TILE_LOG[kind]):
            Z.plan_set(k, digits, cols, kind=plan_kind)
            for op in ("fwd", "coset_inv"):
                if kind == "single":
                    Z.ntt_std_device(ptr(d_in), ptr(d_out), k, Z.OPS[op])
                else:
                    b = Z.Batch([Z.make_job(ptr(d_in), ptr(d_out), k, Z.OPS[op])])
                    b.launch()
                    torch.cuda.synchronize()
                    b.close()
                torch.cuda.synchronize()
                if not torch.equal(d_out, ref[op]):
                    bad.append((op, digits, cols))
    finally:
        Z.plan_reset(k, kind=plan_kind)
        Z.lib().zkgpu_drop_tables(0)
    assert not bad, bad


@pytest.mark.parametrize("k", list(range(12, 24)))
def test_tuned_plans_vs_oracle(torch_cuda, k):
    """The split each plan file holds for 2^k (whatever the tuner chose):
    the single-transform table through the pass kernels, and the batch table
    through the persistent kernel (a one-job Batch always runs it; so does a
    server iteration holding two jobs), against the oracle."""
    torch = torch_cuda
    n = 1 << k
    x = O.gen(1300 + k, n)
    d_in = to_dev(torch, x)
    d_out = dev_empty(torch, n)
    for op in ("fwd", "coset_inv"):
        want = to_dev(torch, O.run_op(op, x, k, threads=8))
        Z.ntt_std_device(ptr(d_in), ptr(d_out), k, Z.OPS[op])
        torch.cuda.synchronize()
        assert torch.equal(d_out, want), ("single", op, Z.plan_get(k, Z.PLAN_SINGLE))
        if k <= 22:
            d_out.zero_()
            b = Z.Batch([Z.make_job(ptr(d_in), ptr(d_out), k, Z.OPS[op])])
            b.launch()
            torch.cuda.synchronize()
            b.close()
            assert torch.equal(d_out, want), ("batch", op, Z.plan_get(k, Z.PLAN_BATCH))
            # two jobs submitted together share the server's iterations
            d_out2 = dev_empty(torch, n)
            d_out.zero_()
            srv = Z.Server()
            try:
                t1 = srv.submit(Z.make_job(ptr(d_in), ptr(d_out), k, Z.OPS[op]))
                t2 = srv.submit(Z.make_job(ptr(d_in), ptr(d_out2), k, Z.OPS[op]))
                r1, r2 = srv.wait(t1), srv.wait(t2)
            finally:
                srv.close()
            assert torch.equal(d_out, want) and torch.equal(d_out2, want), ("served", op)
            assert max(r1.iter_jobs_max, r2.iter_jobs_max) >= 1
