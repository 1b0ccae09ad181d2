"""The C-ABI library without a GPU: it loads, exports every entry point that
include/zkgpu_ntt.h declares, and its host-only logic (domain constants,
pass plans, argument checks, error strings) is right."""
import ctypes
import os
import re

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from zk_gpu_amd.autotune import TILE_LOG, candidates

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "zkgpu_ntt.h")
R = O.R_MOD


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported_and_bound():
    names = _declared()
    assert "parallel_fft_api" in names and "zkgpu_server_submit" in names
    L = ctypes.CDLL(Z.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
        assert n in Z.SIGNATURES, n  # the Python mirror binds it
    assert set(Z.SIGNATURES) == set(names)


def test_drop_in_symbols_match_reference_ffi():
    """src/api.rs:3-13 declares exactly these five entry points."""
    for n in ("do_fft_ffi", "parallel_fft_api", "parallel_ifft_api", "parallel_coeff_to_extended_part_api",
              "parallel_msm_api"):
        assert n in _declared()


def test_version_and_error_strings():
    assert Z.version() >= 100
    for code in range(9):
        s = Z.lib().zkgpu_error_string(code)
        assert s and s != b"unknown error"
    assert Z.lib().zkgpu_error_string(99) == b"unknown error"


@pytest.mark.parametrize("k", [1, 2, 10, 16, 20, 24, 28])
def test_domain_constants(k):
    w = O.root_of_unity(k)
    assert (Z.domain_constant(k, 0) == w.reshape(4)).all()
    assert (Z.domain_constant(k, 1) == O.inv(w).reshape(4)).all()
    assert (Z.domain_constant(k, 2) == O.elem(pow(1 << k, -1, R)).reshape(4)).all()
    assert (Z.domain_constant(k, 3) == O.elem(7).reshape(4)).all()
    assert (Z.domain_constant(k, 4) == O.inv(O.elem(7)).reshape(4)).all()


def test_domain_constant_rejects_bad_arguments():
    out = np.zeros(4, dtype=np.uint64)
    p = out.ctypes.data_as(ctypes.c_void_p)
    assert Z.lib().zkgpu_domain_constant(0, 0, p) == Z.ERR_INVALID
    assert Z.lib().zkgpu_domain_constant(29, 0, p) == Z.ERR_INVALID
    assert Z.lib().zkgpu_domain_constant(10, 5, p) == Z.ERR_INVALID


def _check_plan(k, npass, digits, cols, tile_log):
    assert 1 <= npass <= 3
    assert sum(digits) == k
    for i, (d, c) in enumerate(zip(digits, cols)):
        assert 1 <= d <= 12 and d + c <= tile_log
    log_m = k
    for i in range(npass - 1):
        assert cols[i] <= log_m - digits[i]
        log_m -= digits[i]
    if npass > 1:
        assert cols[-1] <= digits[0]


@pytest.mark.parametrize("kind", [Z.PLAN_SINGLE, Z.PLAN_BATCH])
def test_plans_valid_for_every_size(kind):
    tile_log = 12 if kind == Z.PLAN_SINGLE else 11
    for k in range(1, 29):
        npass, d, c = Z.plan_get(k, kind)
        _check_plan(k, npass, d, c, tile_log)
    assert Z.lib().zkgpu_plan_get_kind(0, kind, None, None) == 0
    assert Z.lib().zkgpu_plan_get_kind(29, kind, None, None) == 0


def test_plan_set_and_reset_restores_tuned_plan():
    for kind in (Z.PLAN_SINGLE, Z.PLAN_BATCH):
        k = 18
        tuned = Z.plan_get(k, kind)
        cands = candidates(k, tile_log_max=TILE_LOG["single" if kind == Z.PLAN_SINGLE else "batch"])
        other = next((d, c) for d, c in cands if [len(d), d, c] != [tuned[0], tuned[1], tuned[2]])
        Z.plan_set(k, other[0], other[1], kind=kind)
        assert Z.plan_get(k, kind) == (len(other[0]), list(other[0]), list(other[1]))
        Z.plan_reset(k, kind=kind)
        assert Z.plan_get(k, kind) == tuned


def test_every_autotune_candidate_is_accepted():
    for kind, name in ((Z.PLAN_SINGLE, "single"), (Z.PLAN_BATCH, "batch")):
        for k in (12, 16, 20, 24):
            for d, c in candidates(k, tile_log_max=TILE_LOG[name]):
                Z.plan_set(k, d, c, kind=kind)
            Z.plan_reset(k, kind=kind)


def test_inconsistent_plans_rejected():
    bad = [
        (20, [10, 9], [0, 0]),   # digits do not sum to log_n
        (20, [13, 7], [0, 0]),   # digit above 12
        (20, [10, 10], [3, 0]),  # tile 2^13 elements
        (24, [12, 12], [0, 13]), # last pass wider than 2^digit0
    ]
    for k, d, c in bad:
        with pytest.raises(Z.ZkGpuError):
            Z.plan_set(k, d, c)
    with pytest.raises(Z.ZkGpuError):
        Z.plan_set(20, [9, 11], [0, 1], kind=Z.PLAN_BATCH)  # 2^12-element tile in a batch


def test_plans_over_workgroup_threads_rejected():
    """A tile of 1- or 2-bit digits keeps fewer than 8 elements per thread, so
    its thread count can pass the workgroup size even when its element count
    fits: 2^1 x 2^10 columns needs 1024 threads (512 per single-pass tile,
    256 per batch tile). Such splits are refused, not launched."""
    with pytest.raises(Z.ZkGpuError):
        Z.plan_set(12, [1, 1, 10], [10, 0, 0], kind=Z.PLAN_BATCH)
    with pytest.raises(Z.ZkGpuError):
        Z.plan_set(13, [1, 12], [10, 0], kind=Z.PLAN_SINGLE)
    with pytest.raises(Z.ZkGpuError):
        Z.plan_set(13, [2, 11], [9, 0], kind=Z.PLAN_BATCH)
    Z.plan_set(13, [2, 11], [9, 0], kind=Z.PLAN_SINGLE)  # 512 threads: allowed for single-pass tiles
    Z.plan_reset(13, kind=Z.PLAN_SINGLE)


def test_drop_in_size_checks_before_any_device_work():
    req = np.zeros((1 << 10, 4), dtype=np.uint64)
    w = O.root_of_unity(10)
    L = Z.lib()
    p = req.ctypes.data_as(ctypes.c_void_p)
    wp = w.ctypes.data_as(ctypes.c_void_p)
    assert L.parallel_fft_api(0, p, wp, None, 1000, 10, 0) == Z.ERR_INVALID       # n != 2^log_n
    assert L.parallel_fft_api(0, p, wp, None, 1 << 29, 29, 0) == Z.ERR_INVALID    # log_n too large
    assert L.parallel_ifft_api(0, p, wp, None, 1 << 10, 10, 0, None) == Z.ERR_INVALID  # no divisor
    assert L.parallel_msm_api(0, None, None, 0, None) == Z.ERR_INVALID  # no result pointer
    assert L.zkgpu_ntt_std_device(0, None, None, 0, 0, None) == Z.ERR_INVALID


def test_job_struct_layout():
    """zkgpu_job as the header lays it out (C and Python sides agree)."""
    assert ctypes.sizeof(Z.Job) == 8 + 8 + 4 * 4 + 3 * 32
    assert Z.Job.omega.offset == 32
    assert ctypes.sizeof(Z.ServerResult) == 4 + 4 + 8 + 4 * 8 + 4 + 4
    assert ctypes.sizeof(Z.ServerStats) == 11 * 8


def test_msm_plan_widths_and_workspace():
    """The MSM plan is host arithmetic: window widths by size (a measured
    table, 15 from 2^14 points, 16 from 2^20), workspace growing with the length, nothing
    for 0 or more than 2^28 points; above 2^27 points the workspace is that
    of one 2^27-point piece."""
    f = Z.lib().zkgpu_msm_workspace_bytes
    c = ctypes.c_uint32(0)
    want = {1: 4, 1 << 7: 5, 1 << 8: 8, 1 << 12: 10, 1 << 13: 11, 1 << 14: 15, 1 << 16: 15, 1 << 19: 15,
            1 << 20: 16, 1 << 24: 16, 1 << 28: 16}
    prev = 0
    for n, cw in sorted(want.items()):
        ws = f(n, ctypes.byref(c))
        assert c.value == cw, n
        assert ws > prev, n
        prev = ws
    assert f(1 << 28, None) == f(1 << 27, None)
    assert f((1 << 27) + 1, None) <= f(1 << 27, None)
    assert f(0, None) == 0
    assert f((1 << 28) + 1, None) == 0
