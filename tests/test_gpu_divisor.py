"""The inverse's divisor. With a multi-pass plan it is folded into the
twiddles of the last pass boundary (applied to every element on load of the
last pass); a one-pass inverse scales its outputs. Any caller divisor, not
only n^-1, gives the oracle's result (reference cuda/fft.cu:153-155 scales
every output by the divisor it is given), and a table built for one divisor
is never reused for another."""
import ctypes

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu

# one-pass (k <= 12), two-pass and three-pass plans
SIZES = [10, 14, 20, 22]


def _divisors(k):
    n_inv = pow(1 << k, -1, O.R_MOD)
    return [O.elem(5 * n_inv % O.R_MOD), O.elem(0x1234567 + k), O.elem(O.R_MOD - 1), O.elem(0)]


@pytest.mark.parametrize("k", SIZES)
def test_inverse_any_divisor_device(torch_cuda, k):
    torch = torch_cuda
    x = O.gen(900 + k, 1 << k)
    w_inv = O.inv(O.root_of_unity(k))
    d_in = to_dev(torch, x)
    for dv in _divisors(k):  # back to back on one size: each needs its own tables
        d_out = dev_empty(torch, 1 << k)
        Z.ntt_device(ptr(d_in), ptr(d_out), k, Z.OP_INV, w_inv, dv)
        torch.cuda.synchronize()
        assert (to_host(d_out) == O.ifft(x, k, omega_inv=w_inv, divisor=dv, threads=8)).all(), (k, O.value(dv[0]))
    # the standard divisor again, after the others
    d_out = dev_empty(torch, 1 << k)
    Z.ntt_std_device(ptr(d_in), ptr(d_out), k, Z.OP_INV)
    torch.cuda.synchronize()
    assert (to_host(d_out) == O.ifft(x, k, threads=8)).all()


@pytest.mark.parametrize("k", [14, 20])
def test_inverse_any_divisor_host_api(k):
    x = O.gen(950 + k, 1 << k)
    w_inv = O.inv(O.root_of_unity(k))
    for dv in _divisors(k)[:2]:
        a = x.copy()
        assert Z.parallel_ifft_api(0, a, w_inv, None, 1 << k, k, 0, dv) == 0
        assert (a == O.ifft(x, k, omega_inv=w_inv, divisor=dv, threads=8)).all()


@pytest.mark.parametrize("k", [12, 16, 20])
def test_inverse_any_divisor_noalloc(torch_cuda, k):
    """The allocation-free entry point rebuilds its scratch tables when only
    the divisor changes."""
    torch = torch_cuda
    x = O.gen(970 + k, 1 << k)
    w_inv = O.inv(O.root_of_unity(k))
    d_in = to_dev(torch, x)
    for dv in _divisors(k)[:3]:
        d_out = dev_empty(torch, 1 << k)
        om, dvv = np.ascontiguousarray(w_inv.reshape(4)), np.ascontiguousarray(dv.reshape(4))
        st = Z.lib().zkgpu_ntt_device_noalloc(0, ptr(d_in), ptr(d_out), k, Z.OP_INV,
                                              om.ctypes.data_as(ctypes.c_void_p), dvv.ctypes.data_as(ctypes.c_void_p),
                                              None, None)
        Z.check(st, "zkgpu_ntt_device_noalloc")
        torch.cuda.synchronize()
        assert (to_host(d_out) == O.ifft(x, k, omega_inv=w_inv, divisor=dv, threads=8)).all(), (k, O.value(dv[0]))


def test_coset_inverse_unaffected_by_folded_tables(torch_cuda):
    """The coset inverse shares the inverse's omega^-1 twiddles but keeps its
    divisor in the post-scale table: it must not pick up the folded ones."""
    torch = torch_cuda
    k = 20
    x = O.gen(990, 1 << k)
    d_in = to_dev(torch, x)
    for op in ("inv", "coset_inv", "inv", "coset_inv"):
        d_out = dev_empty(torch, 1 << k)
        Z.ntt_std_device(ptr(d_in), ptr(d_out), k, Z.OPS[op])
        torch.cuda.synchronize()
        assert (to_host(d_out) == O.run_op(op, x, k, threads=8)).all(), op
