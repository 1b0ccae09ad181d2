"""Parity of the HIP path (through the C-ABI) with the oracle and the
golden fixtures. Bit-exact: the transform is exact field arithmetic."""
import json
import os

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
OPS = ("fwd", "inv", "coset_fwd", "coset_inv")


def test_fill_random_matches_oracle(torch_cuda):
    torch = torch_cuda
    n = 5000
    d = dev_empty(torch, n)
    Z.fill_random(ptr(d), n, 1234, start=77)
    torch.cuda.synchronize()
    assert (to_host(d) == O.gen(1234, n, start=77)).all()


@pytest.mark.parametrize("k", list(range(1, 13)))
def test_std_ops_full_fixture(torch_cuda, k):
    torch = torch_cuda
    f = np.load(os.path.join(GOLD, f"ntt_full_k{k}.npz"))
    x = O.gen(int(f["seed"][0]), 1 << k)
    for op in OPS:
        d_in = to_dev(torch, x)
        d_out = dev_empty(torch, 1 << k)
        Z.ntt_std_device(ptr(d_in), ptr(d_out), k, Z.OPS[op])
        torch.cuda.synchronize()
        got = to_host(d_out)
        assert (got == f[op]).all(), (k, op)
        assert (to_host(d_in) == x).all(), "input modified by an out-of-place call"


@pytest.mark.parametrize("k", [13, 14, 16, 18, 20])
def test_std_ops_vs_oracle(torch_cuda, k):
    torch = torch_cuda
    x = O.gen(500 + k, 1 << k)
    for op in OPS:
        d = to_dev(torch, x)
        Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])  # in place
        torch.cuda.synchronize()
        exp = O.run_op(op, x, k, threads=8)
        assert (to_host(d) == exp).all(), (k, op)


@pytest.mark.parametrize("k", [16, 20, 24])
def test_sampled_fixture(torch_cuda, k):
    torch = torch_cuda
    e = json.load(open(os.path.join(GOLD, "ntt_sampled.json")))[str(k)]
    n = 1 << k
    idx = e["indices"]
    for op in OPS:
        d = dev_empty(torch, n)
        Z.fill_random(ptr(d), n, e["seed"])
        Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])
        torch.cuda.synchronize()
        got = to_host(d)[idx]
        exp = np.array([[int(v, 16) for v in w] for w in e["ops"][op]], dtype=np.uint64)
        assert (got == exp).all(), (k, op)


@pytest.mark.parametrize("k", [10, 15, 17])
def test_drop_in_host_api(torch_cuda, k):
    n = 1 << k
    x = O.gen(900 + k, n)
    w = O.root_of_unity(k)
    winv = O.inv(w)
    div = O.elem(pow(n, -1, O.R_MOD))
    c = O.elem(5)
    tw = O.twiddle_lut(w, k)
    a = x.copy()
    assert Z.parallel_fft_api(0, a, w, tw, n, k, tw.shape[0]) == 0
    assert (a == O.fft(x, k, omega=w, threads=8)).all()
    a = x.copy()
    assert Z.parallel_ifft_api(0, a, winv, None, n, k, 0, div) == 0
    assert (a == O.ifft(x, k, omega_inv=winv, divisor=div, threads=8)).all()
    a = x.copy()
    assert Z.parallel_coeff_to_extended_part_api(0, a, w, tw, n, k, tw.shape[0], c) == 0
    assert (a == O.coset_fft(x, k, c, omega=w, threads=8)).all()


@pytest.mark.parametrize("k", [20, 24])
def test_round_trip_large(torch_cuda, k):
    torch = torch_cuda
    n = 1 << k
    a = dev_empty(torch, n)
    Z.fill_random(ptr(a), n, 42)
    b = dev_empty(torch, n)
    for fwd, inv in ((Z.OP_FWD, Z.OP_INV), (Z.OP_COSET_FWD, Z.OP_COSET_INV)):
        Z.ntt_std_device(ptr(a), ptr(b), k, fwd)
        Z.ntt_std_device(ptr(b), ptr(b), k, inv)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (k, fwd)


@pytest.mark.parametrize("op", OPS)
def test_full_vector_k24_vs_oracle(torch_cuda, op):
    """The 2^24 configuration: every one of the 2^24 outputs of each op equals
    the oracle on the same seeded input."""
    torch = torch_cuda
    k = 24
    n = 1 << k
    x = O.gen(2424, n)
    d = to_dev(torch, x)
    Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])
    torch.cuda.synchronize()
    got = to_host(d)
    del d
    want = O.run_op(op, x, k, threads=16)
    mism = np.nonzero((got != want).any(axis=1))[0]
    assert mism.size == 0, (op, mism[:8])
