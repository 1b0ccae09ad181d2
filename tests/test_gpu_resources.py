"""Resources of the drop-in and device-resident paths: host-buffer calls share
a fixed pool of contexts under an HBM budget (any number of caller threads),
concurrent callers on one stream never mix their passes on its workspace, a
failed table build leaves nothing behind in the resident cache, and opt-in
input validation rejects non-canonical elements."""
import threading

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import dev_empty, ptr, to_dev, to_host

pytestmark = pytest.mark.gpu
R = O.R_MOD


def test_short_lived_threads_bounded_memory(torch_cuda):
    """64 threads, each calling parallel_fft_api once at 2^20 and exiting: the
    device memory the library keeps afterwards stays within its pool of host
    contexts (4 x (32 + 36) bytes per element at 2^20), and every result is
    right."""
    torch = torch_cuda
    k = 20
    n = 1 << k
    w = Z.domain_constant(k, 0)
    inputs = [O.gen(3100 + i, n) for i in range(8)]
    want = []
    for x in inputs:
        d = to_dev(torch, x)
        Z.ntt_std_device(ptr(d), ptr(d), k, Z.OP_FWD)
        torch.cuda.synchronize()
        want.append(to_host(d))
        del d
    torch.cuda.synchronize()
    Z.lib().zkgpu_set_host_budget(0, 1)  # idle contexts drop what earlier calls left them
    Z.lib().zkgpu_set_host_budget(0, 0)
    assert Z.lib().zkgpu_host_held_bytes(0, None) == 0
    a = inputs[0].copy()
    Z.check(Z.parallel_fft_api(0, a, w, None, n, k, 0))  # tables and one context warm
    free_before, _ = torch.cuda.mem_get_info()
    errs = []

    def one_call(i):
        b = inputs[i % 8].copy()
        rc = Z.parallel_fft_api(0, b, w, None, n, k, 0)
        if rc != 0 or not (b == want[i % 8]).all():
            errs.append((i, rc))

    for wave in range(4):  # 16 threads at a time, each started and ended once
        ts = [threading.Thread(target=one_call, args=(16 * wave + i,)) for i in range(16)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    assert not errs, errs
    free_after, _ = torch.cuda.mem_get_info()
    per_ctx = (32 + 36) * n
    held = Z.lib().zkgpu_host_held_bytes(0, None)
    assert held <= 4 * per_ctx, held
    assert free_before - free_after <= 4 * per_ctx + (64 << 20), (free_before, free_after)


def test_host_budget_serialises_calls(torch_cuda):
    """With a host budget that fits two calls' buffers, eight threads calling
    at once never hold more than the budget (the others wait, as the
    reference's GpuLock makes calls wait for GPU memory); results stay right."""
    import ctypes
    k = 18
    n = 1 << k
    w = Z.domain_constant(k, 0)
    x = O.gen(3300, n)
    want = O.fft(x, k, threads=8)
    per_call = (32 + 36) * n
    budget = 2 * per_call + per_call // 2
    prev = Z.lib().zkgpu_set_host_budget(0, budget)
    errs = []

    def call():
        for _ in range(3):
            b = x.copy()
            rc = Z.parallel_fft_api(0, b, w, None, n, k, 0)
            if rc != 0 or not (b == want).all():
                errs.append(rc)

    try:
        ts = [threading.Thread(target=call) for _ in range(8)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        peak = ctypes.c_uint64(0)
        held = Z.lib().zkgpu_host_held_bytes(0, ctypes.byref(peak))
    finally:
        Z.lib().zkgpu_set_host_budget(0, 0)
    assert not errs, errs
    assert held <= budget and peak.value <= budget, (held, peak.value, budget)
    assert prev >= budget


def test_shared_stream_concurrent_device_calls(torch_cuda):
    """Four threads enqueue multi-pass transforms of different sizes on the
    same (null) stream at once: each job's passes use that stream's planar
    workspace, which a lease keeps for one job at a time (and grows between
    jobs), so every output equals the job run alone."""
    torch = torch_cuda
    sizes = [16, 18, 20, 21]
    ops = ["fwd", "inv", "coset_fwd", "coset_inv"]
    ins = [dev_empty(torch, 1 << k) for k in sizes]
    for i, (k, x) in enumerate(zip(sizes, ins)):
        Z.fill_random(ptr(x), 1 << k, 3400 + i)
    torch.cuda.synchronize()
    want = []
    for k, op, x in zip(sizes, ops, ins):
        o = dev_empty(torch, 1 << k)
        Z.ntt_std_device(ptr(x), ptr(o), k, Z.OPS[op])
        torch.cuda.synchronize()
        want.append(o)
    Z.lib().zkgpu_drop_tables(0)
    outs = [[dev_empty(torch, 1 << k) for _ in range(6)] for k in sizes]
    torch.cuda.synchronize()
    errs = []

    def worker(t):
        try:
            for o in outs[t]:
                Z.ntt_std_device(ptr(ins[t]), ptr(o), sizes[t], Z.OPS[ops[t]], stream=None)
        except Exception as e:  # reported below
            errs.append(repr(e))

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    torch.cuda.synchronize()
    assert not errs, errs
    for t in range(4):
        for o in outs[t]:
            assert torch.equal(o, want[t]), (sizes[t], ops[t])


def test_failed_table_build_publishes_nothing(torch_cuda):
    """A table build that fails half way (an allocation refused by the test
    hook after the first table of the job was allocated) returns an error and
    leaves no table in the resident cache; the next call builds them all and
    is right."""
    torch = torch_cuda
    k = 17
    n = 1 << k
    x = O.gen(3500, n)
    c = O.elem(11)
    w = Z.domain_constant(k, 0)
    Z.lib().zkgpu_drop_tables(0)
    before = Z.lib().zkgpu_table_bytes(0)
    d_in = to_dev(torch, x)
    d_out = dev_empty(torch, n)
    torch.cuda.synchronize()
    for skip in (0, 1, 2, 3):  # fail the 1st, 2nd, ... allocation of the build
        Z.lib().zkgpu_debug_fail_allocs(skip, 1)
        try:
            with pytest.raises(Z.ZkGpuError) as ei:
                Z.ntt_device(ptr(d_in), ptr(d_out), k, Z.OP_COSET_FWD, w, None, c)
            assert ei.value.code == Z.ERR_NOMEM
        finally:
            Z.lib().zkgpu_debug_fail_allocs(0, 0)
        torch.cuda.synchronize()
        assert Z.lib().zkgpu_table_bytes(0) == before, skip
    Z.ntt_device(ptr(d_in), ptr(d_out), k, Z.OP_COSET_FWD, w, None, c)
    torch.cuda.synchronize()
    assert (to_host(d_out) == O.coset_fft(x, k, c, threads=8)).all()
    assert Z.lib().zkgpu_table_bytes(0) > before


def test_validation_rejects_noncanonical_input(torch_cuda):
    k = 12
    n = 1 << k
    w = Z.domain_constant(k, 0)
    x = O.gen(3600, n)
    prev = Z.lib().zkgpu_set_validation(1)
    try:
        bad = x.copy()
        bad[777] = O.to_limbs(R)  # == r: not canonical
        keep = bad.copy()
        assert Z.parallel_fft_api(0, bad, w, None, n, k, 0) == Z.ERR_NONCANONICAL
        assert (bad == keep).all()  # the caller's buffer is left as it was
        bad[777] = O.to_limbs((1 << 256) - 1)
        assert Z.parallel_fft_api(0, bad, w, None, n, k, 0) == Z.ERR_NONCANONICAL
        ok = x.copy()
        ok[777] = O.to_limbs(R - 1)
        want = O.fft(ok, k, threads=4)
        assert Z.parallel_fft_api(0, ok, w, None, n, k, 0) == 0
        assert (ok == want).all()
    finally:
        Z.lib().zkgpu_set_validation(prev)
    # off: no check (the reference does none)
    assert Z.lib().zkgpu_set_validation(0) == 0


def test_do_fft_ffi_smoke(torch_cuda):
    """The reference's device smoke entry point (src/api.rs:4): 0 on a
    working device, also from several threads at once."""
    assert Z.do_fft_ffi(0) == 0
    rcs = []
    ts = [threading.Thread(target=lambda: rcs.append(Z.do_fft_ffi(0))) for _ in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert rcs == [0, 0, 0, 0]
    assert Z.do_fft_ffi(7) == Z.ERR_INVALID  # no such device


@pytest.mark.parametrize("k", [10, 16, 20])
def test_coset_ifft_api_vs_oracle(torch_cuda, k):
    """zkgpu_coset_ifft_api (host buffers) against the oracle's coset inverse,
    and as the inverse of parallel_coeff_to_extended_part_api."""
    n = 1 << k
    x = O.gen(3700 + k, n)
    winv = Z.domain_constant(k, 1)
    ninv = Z.domain_constant(k, 2)
    c = O.elem(5)
    cinv = O.inv(c)
    a = x.copy()
    Z.check(Z.zkgpu_coset_ifft_api(0, a, winv, n, k, ninv, cinv))
    assert (a == O.coset_ifft(x, k, c, threads=8)).all()
    b = x.copy()
    Z.check(Z.parallel_coeff_to_extended_part_api(0, b, Z.domain_constant(k, 0), None, n, k, 0, c))
    Z.check(Z.zkgpu_coset_ifft_api(0, b, winv, n, k, ninv, cinv))
    assert (b == x).all()
