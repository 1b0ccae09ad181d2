"""Serving (continuous batching): every job's result must be bit-identical to
the same job run alone, whatever else shares its iterations, and the
admission budget must hold."""
import random
import time

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_host

pytestmark = pytest.mark.gpu
OPNAMES = ("fwd", "inv", "coset_fwd", "coset_inv")


def _alone(torch, x, k, op):
    out = dev_empty(torch, 1 << k)
    Z.ntt_std_device(ptr(x), ptr(out), k, Z.OPS[op])
    torch.cuda.synchronize()
    return to_host(out)


def _server_out(torch, r, k):
    """Copy a server-owned result (device pointer) to the host."""
    n = 1 << k
    t = torch.empty(n * 4, dtype=torch.int64, device="cuda")
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipMemcpy(ctypes.c_void_p(ptr(t)), ctypes.c_void_p(r.d_out), ctypes.c_size_t(n * 32), 3) == 0
    torch.cuda.synchronize()
    return to_host(t)


def test_serve_alone_vs_mix(torch_cuda):
    torch = torch_cuda
    rnd = random.Random(11)
    sizes = [rnd.randint(10, 20) for _ in range(48)] + [1, 3, 9, 21]
    ins, outs, meta = [], [], []
    for i, k in enumerate(sizes):
        x = dev_empty(torch, 1 << k)
        Z.fill_random(ptr(x), 1 << k, 50_000 + i)
        ins.append(x)
        own = i % 3 == 0  # every third job asks for a server-owned result
        outs.append(None if own else dev_empty(torch, 1 << k))
        meta.append((k, rnd.choice(OPNAMES), 50_000 + i))
    torch.cuda.synchronize()
    srv = Z.Server()
    try:
        tickets = []
        for i, ((k, op, _), x, o) in enumerate(zip(meta, ins, outs)):
            tickets.append(srv.submit(Z.make_job(ptr(x), ptr(o) if o is not None else None, k, Z.OPS[op])))
            if i % 5 == 4:
                time.sleep(0.0005 * rnd.random())  # later jobs arrive while earlier ones run
        for t, (k, op, seed), x, o in zip(tickets, meta, ins, outs):
            r = srv.wait(t)
            assert r.done == 1
            got = to_host(o) if o is not None else _server_out(torch, r, k)
            assert (got == _alone(torch, x, k, op)).all(), (k, op)
            if k <= 12:
                assert (got == O.run_op(op, O.gen(seed, 1 << k), k, threads=4)).all(), (k, op)
            assert r.t_submit_ns <= r.t_admit_ns <= r.t_start_ns <= r.t_done_ns
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == len(sizes) and st["failed"] == 0
        assert st["iterations"] < sum(Z.plan_get(k, Z.PLAN_BATCH)[0] for k in sizes)  # jobs shared iterations
        assert st["in_use_bytes"] == 0
    finally:
        srv.close()


def test_serve_budget_admission(torch_cuda):
    torch = torch_cuda
    k = 18
    n = 1 << k
    per_job = 32 * n + 36 * n  # server-owned output + workspace
    srv = Z.Server(budget_bytes=2 * per_job + 1024, max_active=8)
    try:
        x = dev_empty(torch, n)
        Z.fill_random(ptr(x), n, 77)
        torch.cuda.synchronize()
        ref = _alone(torch, x, k, "coset_fwd")
        tickets = [srv.submit(Z.make_job(ptr(x), None, k, Z.OP_COSET_FWD)) for _ in range(6)]
        for t in tickets:
            r = srv.wait(t)
            assert (_server_out(torch, r, k) == ref).all()
            srv.release(t)  # frees budget for the jobs still pending
        st = srv.stats()
        assert st["peak_in_use_bytes"] <= st["budget_bytes"]
        assert st["completed"] == 6
        with pytest.raises(Z.ZkGpuError):
            srv.submit(Z.make_job(ptr(x), None, 20, Z.OP_FWD))  # footprint alone exceeds the budget
    finally:
        srv.close()


def test_serve_destroy_with_jobs_in_flight(torch_cuda):
    torch = torch_cuda
    srv = Z.Server()
    xs = []
    for i in range(8):
        x = dev_empty(torch, 1 << 16)
        Z.fill_random(ptr(x), 1 << 16, 90 + i)
        xs.append(x)
    torch.cuda.synchronize()
    for x in xs:
        srv.submit(Z.make_job(ptr(x), ptr(x), 16, Z.OP_FWD))
    srv.close()  # admitted jobs finish, pending ones are cancelled; must not hang
    torch.cuda.synchronize()


def test_serve_concurrent_submitters(torch_cuda):
    """Four host threads submit and wait at once; every result equals the job
    run alone."""
    import threading
    torch = torch_cuda
    rnd = random.Random(21)
    specs = [[(rnd.randint(10, 17), rnd.choice(OPNAMES)) for _ in range(10)] for _ in range(4)]
    ins = [[dev_empty(torch, 1 << k) for k, _ in sp] for sp in specs]
    for t, sp in enumerate(specs):
        for i, (k, _) in enumerate(sp):
            Z.fill_random(ptr(ins[t][i]), 1 << k, 60_000 + 100 * t + i)
    outs = [[dev_empty(torch, 1 << k) for k, _ in sp] for sp in specs]
    torch.cuda.synchronize()
    srv = Z.Server()
    errs = []

    def worker(t):
        try:
            tickets = [srv.submit(Z.make_job(ptr(x), ptr(o), k, Z.OPS[op]))
                       for (k, op), x, o in zip(specs[t], ins[t], outs[t])]
            for tk in tickets:
                srv.wait(tk)
                srv.release(tk)
        except Exception as e:  # reported below
            errs.append((t, repr(e)))

    try:
        ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        assert not errs, errs
        for t, sp in enumerate(specs):
            for (k, op), x, o in zip(sp, ins[t], outs[t]):
                assert (to_host(o) == _alone(torch, x, k, op)).all(), (t, k, op)
        assert srv.stats()["completed"] == 40
    finally:
        srv.close()
