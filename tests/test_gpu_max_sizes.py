"""The largest transforms the library accepts (2^26 and 2^28 elements: byte
offsets into the 36-byte workspace pass 2^32). Size-independent properties
(round trips) and, at 2^26, outputs checked against the definition
X_j = a(omega^j) at sampled indices."""
import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_host

pytestmark = pytest.mark.gpu
R = O.R_MOD


def test_k26_definition_and_round_trip(torch_cuda):
    torch = torch_cuda
    k = 26
    n = 1 << k
    x = dev_empty(torch, n)
    Z.fill_random(ptr(x), n, 2626)
    y = dev_empty(torch, n)
    Z.ntt_std_device(ptr(x), ptr(y), k, Z.OP_FWD)
    torch.cuda.synchronize()
    a = to_host(x)
    X = y.view(-1, 4)
    w = O.value(O.root_of_unity(k)[0])
    for j in (1, n // 2 + 12345, n - 1):
        want = O.evaluate(a, O.elem(pow(w, j, R)))
        got = X[j].cpu().numpy().view(np.uint64).reshape(1, 4)
        assert (got == want).all(), j
    Z.ntt_std_device(ptr(y), ptr(y), k, Z.OP_INV)
    torch.cuda.synchronize()
    assert torch.equal(x, y)
    del a, X, y
    Z.lib().zkgpu_drop_tables(0)


def test_k28_round_trips(torch_cuda):
    torch = torch_cuda
    k = 28
    n = 1 << k
    x = dev_empty(torch, n)
    Z.fill_random(ptr(x), n, 2828)
    y = dev_empty(torch, n)
    for fwd, inv in ((Z.OP_FWD, Z.OP_INV), (Z.OP_COSET_FWD, Z.OP_COSET_INV)):
        Z.ntt_std_device(ptr(x), ptr(y), k, fwd)
        Z.ntt_std_device(ptr(y), ptr(y), k, inv)
        torch.cuda.synchronize()
        assert torch.equal(x, y), (fwd, inv)
        Z.lib().zkgpu_drop_tables(0)
    del x, y
    torch.cuda.empty_cache()
