"""Batched transforms (one persistent launch per pass level, replayed from a
captured graph): every job must equal the same job run alone, bit for bit,
and the oracle where the oracle is cheap."""
import random

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_host

pytestmark = pytest.mark.gpu
OPNAMES = ("fwd", "inv", "coset_fwd", "coset_inv")


def _mixed_jobs(torch, sizes, seed, in_place_every=3):
    rnd = random.Random(seed)
    ins, outs, jobs, meta = [], [], [], []
    for i, k in enumerate(sizes):
        n = 1 << k
        a = dev_empty(torch, n)
        Z.fill_random(ptr(a), n, 10_000 + i)
        op = rnd.choice(OPNAMES)
        b = a if i % in_place_every == 0 else dev_empty(torch, n)
        ins.append(a)
        outs.append(b)
        jobs.append(Z.make_job(ptr(a), ptr(b), k, Z.OPS[op]))
        meta.append((k, op, 10_000 + i, b is a))
    return ins, outs, jobs, meta


def _alone(torch, k, op, seed):
    n = 1 << k
    a = dev_empty(torch, n)
    Z.fill_random(ptr(a), n, seed)
    Z.ntt_std_device(ptr(a), ptr(a), k, Z.OPS[op])
    torch.cuda.synchronize()
    return to_host(a)


def test_batch_mixed_sizes_match_alone(torch_cuda):
    torch = torch_cuda
    sizes = [1, 2, 3, 5, 8, 10, 11, 12, 13, 14, 16, 17, 18, 19, 20, 12, 9, 15, 11, 16]
    ins, outs, jobs, meta = _mixed_jobs(torch, sizes, seed=1)
    b = Z.Batch(jobs)
    info = b.info()
    assert info["elements"] == sum(1 << k for k in sizes)
    b.launch()
    torch.cuda.synchronize()
    for (k, op, seed, inplace), out in zip(meta, outs):
        got = to_host(out)
        assert (got == _alone(torch, k, op, seed)).all(), (k, op)
        if k <= 12:
            x = O.gen(seed, 1 << k)
            assert (got == O.run_op(op, x, k, threads=4)).all(), (k, op)
    b.close()


def test_batch_graph_replays_identical(torch_cuda):
    torch = torch_cuda
    sizes = [12, 14, 16, 13, 15]
    rnd = random.Random(5)
    ins, outs, jobs = [], [], []
    for i, k in enumerate(sizes):
        n = 1 << k
        a = dev_empty(torch, n)
        Z.fill_random(ptr(a), n, 20_000 + i)
        o = dev_empty(torch, n)
        ins.append(a)
        outs.append(o)
        jobs.append(Z.make_job(ptr(a), ptr(o), k, Z.OPS[rnd.choice(OPNAMES)]))
    b = Z.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    first = [to_host(o) for o in outs]
    for rep in range(4):
        for o in outs:
            o.zero_()
        # unrelated eager work between replays
        scratch = torch.empty(1 << 20, dtype=torch.int64, device="cuda").fill_(rep)
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        for f, o in zip(first, outs):
            assert (to_host(o) == f).all(), rep
        del scratch
    b.close()


def test_batch_round_trip_and_explicit_constants(torch_cuda):
    torch = torch_cuda
    k = 14
    n = 1 << k
    x = dev_empty(torch, n)
    Z.fill_random(ptr(x), n, 77)
    y = dev_empty(torch, n)
    w = Z.domain_constant(k, 0)
    winv = Z.domain_constant(k, 1)
    ninv = Z.domain_constant(k, 2)
    c = O.elem(5).reshape(4)
    cinv = O.inv(O.elem(5)).reshape(4)
    fwd = Z.Batch([Z.make_job(ptr(x), ptr(y), k, Z.OP_COSET_FWD, omega=w, divisor=ninv, c=c)])
    inv = Z.Batch([Z.make_job(ptr(y), ptr(y), k, Z.OP_COSET_INV, omega=winv, divisor=ninv, c=cinv)])
    fwd.launch()
    inv.launch()
    torch.cuda.synchronize()
    assert torch.equal(x, y)
    fwd.close()
    inv.close()


def test_batch_c4_shape(torch_cuda):
    """The batched configuration: 64 polynomials, sizes log-uniform in
    2^12..2^20, ops uniform over the four, in one captured batch. Every one of
    the 64 outputs equals the oracle (halo2 best_fft restated) on the same
    input."""
    torch = torch_cuda
    rnd = random.Random(4)
    sizes = [rnd.randint(12, 20) for _ in range(64)]
    ins, outs, jobs, meta = _mixed_jobs(torch, sizes, seed=4, in_place_every=1)
    b = Z.Batch(jobs)
    assert b.info()["levels"] >= 2
    b.launch()
    torch.cuda.synchronize()
    bad = []
    for idx, ((k, op, seed, _), out) in enumerate(zip(meta, outs)):
        want = O.run_op(op, O.gen(seed, 1 << k), k, threads=16)
        if not (to_host(out) == want).all():
            bad.append((idx, k, op))
    b.close()
    assert not bad, bad
