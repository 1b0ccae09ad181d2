"""The entry point that allocates nothing (zkgpu_ntt_device_noalloc): its
outputs equal zkgpu_ntt_device's bit for bit, on consecutive calls whose
sizes shrink (the scratch keeps the previous call's contents), it makes no
device allocation, and it refuses sizes beyond its scratch."""
import ctypes

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
OPS = ("fwd", "inv", "coset_fwd", "coset_inv")


def _consts(k, op, g=7):
    inverse = op in ("inv", "coset_inv")
    w = Z.domain_constant(k, 1 if inverse else 0)
    d = Z.domain_constant(k, 2)
    c = O.elem(g).reshape(4)
    if op == "coset_inv":
        c = O.inv(O.elem(g)).reshape(4)
    return w, d, c


def _noalloc(d_in, d_out, k, op, g=7):
    w, d, c = _consts(k, op, g)
    return Z.lib().zkgpu_ntt_device_noalloc(0, d_in, d_out, k, Z.OPS[op], Z._ptr(w), Z._ptr(d), Z._ptr(c), None)


def _reference(torch, d_in, k, op, g=7):
    out = dev_empty(torch, 1 << k)
    w, d, c = _consts(k, op, g)
    Z.ntt_device(d_in, ptr(out), k, Z.OPS[op], w, d, c)
    torch.cuda.synchronize()
    return out


def test_scratch_limits():
    mx = ctypes.c_uint32(0)
    assert Z.lib().zkgpu_noalloc_scratch_bytes(ctypes.byref(mx)) == 120 << 20
    assert mx.value == 20


@pytest.mark.parametrize("op", OPS)
def test_noalloc_matches_device_entry_point(torch_cuda, op):
    """Two calls in a row, the second smaller (2^20 then 2^14; 2^13 then 2^9
    single-pass), each equal to zkgpu_ntt_device on the same input."""
    torch = torch_cuda
    for ks in ((20, 14), (13, 9)):
        for i, k in enumerate(ks):
            n = 1 << k
            x = dev_empty(torch, n)
            Z.fill_random(ptr(x), n, 4000 + 10 * k + i)
            y = dev_empty(torch, n)
            torch.cuda.synchronize()
            assert _noalloc(ptr(x), ptr(y), k, op) == 0
            torch.cuda.synchronize()
            assert torch.equal(y, _reference(torch, ptr(x), k, op)), (op, k)


def test_noalloc_reuses_and_rebuilds_tables(torch_cuda):
    """Same size with a new coset (tables rebuilt in scratch), the same call
    twice (tables reused), in place, and one case against the oracle."""
    torch = torch_cuda
    k = 12
    n = 1 << k
    xs = O.gen(4100, n)
    x = to_dev(torch, xs)
    y = dev_empty(torch, n)
    for g in (7, 5, 5, 7):
        assert _noalloc(ptr(x), ptr(y), k, "coset_fwd", g) == 0
        torch.cuda.synchronize()
        assert torch.equal(y, _reference(torch, ptr(x), k, "coset_fwd", g)), g
    assert (to_host(y) == O.coset_fft(xs, k, O.elem(7), threads=4)).all()
    z = x.clone()
    assert _noalloc(ptr(z), ptr(z), k, "inv") == 0  # in place
    torch.cuda.synchronize()
    assert torch.equal(z, _reference(torch, ptr(x), k, "inv"))


def test_noalloc_allocates_nothing(torch_cuda):
    """Calls with constants no resident table has seen leave the device's free
    memory and the resident table cache unchanged."""
    torch = torch_cuda
    k = 18
    n = 1 << k
    x = dev_empty(torch, n)
    Z.fill_random(ptr(x), n, 4200)
    y = dev_empty(torch, n)
    torch.cuda.synchronize()
    assert _noalloc(ptr(x), ptr(y), k, "fwd") == 0  # the library is set up on this device
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    tables0 = Z.lib().zkgpu_table_bytes(0)
    for g in (11, 13, 17):
        for op in ("coset_fwd", "coset_inv"):
            assert _noalloc(ptr(x), ptr(y), k, op, g) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    assert Z.lib().zkgpu_table_bytes(0) == tables0


def test_noalloc_rejects_beyond_scratch(torch_cuda):
    torch = torch_cuda
    x = dev_empty(torch, 1 << 10)
    assert _noalloc(ptr(x), ptr(x), 21, "fwd") == Z.ERR_SCRATCH
    w = Z.domain_constant(10, 0)
    assert Z.lib().zkgpu_ntt_device_noalloc(0, ptr(x), ptr(x), 0, 0, Z._ptr(w), None, None, None) == Z.ERR_INVALID
    assert Z.lib().zkgpu_ntt_device_noalloc(0, ptr(x), ptr(x), 10, 2, Z._ptr(w), None, None, None) == Z.ERR_INVALID
    assert Z.lib().zkgpu_ntt_device_noalloc(0, None, ptr(x), 10, 0, None, None, None, None) == Z.ERR_INVALID
