"""The drop-in host-buffer entry points (pinned, chunked staging of the
caller's memory): results equal the device-resident path and the oracle at
sizes below, at and above the staging chunk, and under concurrent callers."""
import threading

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
R = O.R_MOD


def _device(torch, x, k, op):
    d = to_dev(torch, x)
    Z.ntt_std_device(ptr(d), ptr(d), k, Z.OPS[op])
    torch.cuda.synchronize()
    return to_host(d)


@pytest.mark.parametrize("k", [14, 15, 17, 18, 19, 20, 21])  # direct copy, 1 MiB .. 64 MiB (1..8 chunks)
def test_host_api_matches_device_path(torch_cuda, k):
    torch = torch_cuda
    n = 1 << k
    x = O.gen(900 + k, n)
    w, winv = Z.domain_constant(k, 0), Z.domain_constant(k, 1)
    ninv, g = Z.domain_constant(k, 2), Z.domain_constant(k, 3)
    a = x.copy()
    Z.check(Z.parallel_fft_api(0, a, w, None, n, k, 0))
    assert (a == _device(torch, x, k, "fwd")).all()
    b = a.copy()
    Z.check(Z.parallel_ifft_api(0, b, winv, None, n, k, 0, ninv))
    assert (b == x).all()
    c = x.copy()
    Z.check(Z.parallel_coeff_to_extended_part_api(0, c, w, None, n, k, 0, g))
    assert (c == _device(torch, x, k, "coset_fwd")).all()
    if k <= 16:
        assert (a == O.fft(x, k, threads=4)).all()


def test_host_api_concurrent_callers(torch_cuda):
    k = 20
    n = 1 << k
    w = Z.domain_constant(k, 0)
    inputs = [O.gen(950 + t, n) for t in range(4)]
    want = []
    for x in inputs:
        a = x.copy()
        Z.check(Z.parallel_fft_api(0, a, w, None, n, k, 0))
        want.append(a)
    got = [x.copy() for x in inputs]
    errs = []

    def worker(i):
        for _ in range(3):
            a = inputs[i].copy()
            rc = Z.parallel_fft_api(0, a, w, None, n, k, 0)
            if rc != 0 or not (a == want[i]).all():
                errs.append((i, rc))
        got[i] = a

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for g_, w_ in zip(got, want):
        assert (g_ == w_).all()


@pytest.mark.parametrize("log_n", [8, 12, 16])
def test_reference_caller_protocol(torch_cuda, log_n):
    """The reference's own caller (src/main.rs:50-91, 128-162): elements
    Fr::from_raw([i, 9720557559516996230, 14346186778711300467,
    2065298866258849470]) in Montgomery form, omega derived through the
    extended domain (j = 3 -> extended_k = k + 1, squared back down), the
    sparse twiddle table generate_twiddle_lookup_table(omega, k, 10, true)
    passed along, parallel_fft_api in place. The result equals the oracle and,
    at sampled indices, the definition X_j = a(omega^j)."""
    n = 1 << log_n
    hi = 9720557559516996230 << 64 | 14346186778711300467 << 128 | 2065298866258849470 << 192
    vals = [i | hi for i in range(n)]
    data = np.concatenate([O.elem(v) for v in vals])
    # omega: the 2^28-adic root squared down to the extended domain, then to 2^k
    ext_k = log_n
    while (1 << ext_k) < n * 2:
        ext_k += 1
    w_ext = O.value(O.root_of_unity(28)[0])
    for _ in range(ext_k, 28):
        w_ext = w_ext * w_ext % R
    w = w_ext
    for _ in range(log_n, ext_k):
        w = w * w % R
    omega = O.elem(w)
    assert (omega == O.root_of_unity(log_n)).all()
    twiddle = O.twiddle_lut(omega, log_n)
    a = data.copy()
    rc = Z.lib().parallel_fft_api(0, Z._ptr(a), Z._ptr(omega.reshape(4)), Z._ptr(twiddle), n, log_n,
                                  twiddle.shape[0])
    assert rc == 0
    assert (a == O.fft(data, log_n, threads=4)).all()
    for j in (0, 1, n // 2 + 3, n - 1):
        x = pow(w, j, R)
        acc = 0
        for v in reversed(vals):
            acc = (acc * x + v) % R
        assert O.value(a[j]) == acc, j
