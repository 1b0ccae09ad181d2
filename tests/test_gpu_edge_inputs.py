"""Edge-value inputs through every path (single transforms, batches, the
server) against the oracle. The kernels keep field elements in a redundant
29-bit-limb form whose normalisations are scheduled at compile time from
tracked limb and value bounds; the largest canonical values (r - 1), values
whose top limb equals r's, and structured vectors (impulse, constant,
alternating) are where a wrong bound or a missed reduction would show. The
reference's own add compares only the top limb and can wrap for such values
(cuda/gpu_math.h:255-264)."""
import random

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
R = O.R_MOD
OPS = ("fwd", "inv", "coset_fwd", "coset_inv")
PATTERNS = ("all_rm1", "alt_0_rm1", "impulse_rm1", "constant_top", "top_limb_eq")


def pattern(name, k, seed=0):
    """Montgomery words (n, 4) of an edge-value input of 2^k elements."""
    n = 1 << k
    rm1 = O.to_limbs(R - 1)
    top = R >> 192  # r's top 64-bit word
    if name == "all_rm1":
        return np.tile(rm1, (n, 1))
    if name == "alt_0_rm1":
        a = np.zeros((n, 4), dtype=np.uint64)
        a[1::2] = rm1
        return a
    if name == "impulse_rm1":
        a = np.zeros((n, 4), dtype=np.uint64)
        a[0] = rm1
        a[n // 3] = rm1
        return a
    if name == "constant_top":
        return np.tile(O.to_limbs(R - (1 << 190) - 12345), (n, 1))
    if name == "top_limb_eq":
        # top word equal to r's, lower words random below r's
        rng = np.random.default_rng(1000 + seed)
        a = rng.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, size=(n, 4), dtype=np.uint64)
        a[:, 3] = np.uint64(top)
        low = R & ((1 << 192) - 1)
        a[:, 2] = np.minimum(a[:, 2], np.uint64((low >> 128) - 1))
        return np.ascontiguousarray(a)
    raise ValueError(name)


def _check_canonical(a):
    assert all(O.from_limbs(row) < R for row in a[:: max(1, a.shape[0] // 64)])


@pytest.mark.parametrize("k", [12, 16, 20])
@pytest.mark.parametrize("name", PATTERNS)
def test_single_edge_inputs(torch_cuda, k, name):
    torch = torch_cuda
    x = pattern(name, k, seed=k)
    _check_canonical(x)
    for op in OPS:
        d = to_dev(torch, x)
        Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])
        torch.cuda.synchronize()
        assert (to_host(d) == O.run_op(op, x, k, threads=16)).all(), (k, name, op)


@pytest.mark.parametrize("name", PATTERNS)
def test_single_edge_inputs_k24(torch_cuda, name):
    torch = torch_cuda
    k = 24
    x = pattern(name, k, seed=k)
    for op in ("fwd", "coset_inv"):
        d = to_dev(torch, x)
        Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])
        torch.cuda.synchronize()
        got = to_host(d)
        del d
        assert (got == O.run_op(op, x, k, threads=16)).all(), (name, op)


def test_batch_edge_inputs(torch_cuda):
    """Every pattern at k = 12, 16 and 20 with every op, all in one batch
    (persistent kernel)."""
    torch = torch_cuda
    jobs, outs, meta, keep = [], [], [], []
    for k in (12, 16, 20):
        for pi, name in enumerate(PATTERNS):
            x = pattern(name, k, seed=k + pi)
            op = OPS[(pi + k) % 4]
            d_in = to_dev(torch, x)
            d_out = dev_empty(torch, 1 << k)
            keep.append(d_in)
            outs.append(d_out)
            jobs.append(Z.make_job(ptr(d_in), ptr(d_out), k, Z.OPS[op]))
            meta.append((k, name, op, x))
    b = Z.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    b.close()
    bad = [(k, name, op) for (k, name, op, x), o in zip(meta, outs)
           if not (to_host(o) == O.run_op(op, x, k, threads=16)).all()]
    assert not bad, bad


def test_served_edge_inputs(torch_cuda):
    """The same patterns (k = 12, 16, 20, and 24 for two of them) submitted to
    a server together, so that they share iterations."""
    torch = torch_cuda
    rnd = random.Random(3)
    specs = [(k, name, OPS[rnd.randrange(4)]) for k in (12, 16, 20) for name in PATTERNS]
    specs += [(24, "all_rm1", "inv"), (24, "top_limb_eq", "coset_fwd")]
    ins, outs, xs = [], [], []
    for k, name, op in specs:
        x = pattern(name, k, seed=7 * k)
        xs.append(x)
        ins.append(to_dev(torch, x))
        outs.append(dev_empty(torch, 1 << k))
    torch.cuda.synchronize()
    srv = Z.Server()
    try:
        tickets = [srv.submit(Z.make_job(ptr(i), ptr(o), k, Z.OPS[op])) for (k, _, op), i, o in zip(specs, ins, outs)]
        res = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    assert any(r.iter_jobs_max >= 2 for r in res)
    bad = [(k, name, op) for (k, name, op), x, o in zip(specs, xs, outs)
           if not (to_host(o) == O.run_op(op, x, k, threads=16)).all()]
    assert not bad, bad
