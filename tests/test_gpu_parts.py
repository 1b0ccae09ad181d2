"""Extended-domain parts (m coset-forward transforms of one polynomial in one
batched run): each part equals the oracle's coset FFT with its coset, and the
same transform run alone."""
import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
R = O.R_MOD


def _cosets(k, parts):
    """halo2's extended parts: c_j = g * omega_ext^j, g = 7, omega_ext of order 2^(k + log2 parts)."""
    e = parts.bit_length() - 1
    w_ext = O.value(O.root_of_unity(k + e)[0])
    return np.concatenate([O.elem(7 * pow(w_ext, j, R) % R) for j in range(parts)])


@pytest.mark.parametrize("k,parts", [(10, 2), (12, 4), (14, 8)])
def test_parts_device_vs_oracle(torch_cuda, k, parts):
    torch = torch_cuda
    n = 1 << k
    x = O.gen(700 + k, n)
    cs = _cosets(k, parts)
    d_in = to_dev(torch, x)
    outs = [dev_empty(torch, n) for _ in range(parts)]
    Z.coeff_to_extended_parts_device(ptr(d_in), [ptr(o) for o in outs], k, O.root_of_unity(k), cs)
    for j, o in enumerate(outs):
        assert (to_host(o) == O.coset_fft(x, k, cs[j:j + 1], threads=4)).all(), j
    assert (to_host(d_in) == x).all()  # the input is left as it was


@pytest.mark.parametrize("k,parts", [(17, 8), (19, 4)])  # batched / one transform per part
def test_parts_device_vs_single_calls(torch_cuda, k, parts):
    torch = torch_cuda
    n = 1 << k
    d_in = dev_empty(torch, n)
    Z.fill_random(ptr(d_in), n, 31)
    cs = _cosets(k, parts)
    w = O.root_of_unity(k)
    outs = [dev_empty(torch, n) for _ in range(parts)]
    Z.coeff_to_extended_parts_device(ptr(d_in), [ptr(o) for o in outs], k, w, cs)
    ref = dev_empty(torch, n)
    for j, o in enumerate(outs):
        Z.ntt_device(ptr(d_in), ptr(ref), k, Z.OP_COSET_FWD, w, None, cs[j])
        torch.cuda.synchronize()
        assert torch.equal(o, ref), j


def test_parts_host_api(torch_cuda):
    k, parts = 13, 4
    n = 1 << k
    x = O.gen(801, n)
    cs = _cosets(k, parts)
    got = Z.coeff_to_extended_parts_api(0, x, O.root_of_unity(k), n, k, cs)
    for j in range(parts):
        want = x.copy()
        Z.check(Z.parallel_coeff_to_extended_part_api(0, want, O.root_of_unity(k), None, n, k, 0, cs[j]))
        assert (got[j] == want).all(), j
        assert (got[j] == O.coset_fft(x, k, cs[j:j + 1], threads=4)).all(), j


def test_parts_reject_bad_arguments(torch_cuda):
    torch = torch_cuda
    x = dev_empty(torch, 1 << 10)
    with pytest.raises(Z.ZkGpuError):
        Z.coeff_to_extended_parts_device(ptr(x), [ptr(x)], 10, O.root_of_unity(10), O.elem(7))  # output aliases input
    with pytest.raises(Z.ZkGpuError):
        Z.coeff_to_extended_parts_device(ptr(x), [], 10, O.root_of_unity(10), np.zeros((0, 4), np.uint64))
