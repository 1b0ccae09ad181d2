"""G1 multi-scalar multiplication on the device (parallel_msm_api and
zkgpu_msm_device) against the golden vectors, the oracle (bit-serial Straus
on the CPU, a different summation order), and the closed form
sum_i s_i (a + i d) G = (sum_i s_i (a + i d) mod r) G. The result is a group
element: compared as affine coordinates (the drop-in returns z = 1)."""
import json
import os
import random

import numpy as np
import pytest

import zk_gpu_amd as Z
from oracle import oracle as O
from tests.gpu_util import ptr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "g1_vectors.json")
R = O.R_MOD


def _arr(rows):
    return np.ascontiguousarray(np.array([[int(w, 16) for w in row] for row in rows], dtype=np.uint64))


def _affine(acc):
    """(3, 4) result words -> None (identity) or canonical (x, y); z must be 1 or 0"""
    acc = np.asarray(acc, dtype=np.uint64).reshape(3, 4)
    if not acc.any():
        return None
    assert O.fq_value(acc[2]) == 1  # z = 1 in Montgomery form
    return O.g1_to_affine(acc.reshape(12))


def _weighted_sum(scalars, a, d):
    """sum_i s_i (a + i d) mod r, s_i the canonical values of the Montgomery
    words: R^-1 sum_i m_i (a + i d) with m_i the stored integers"""
    w = scalars.astype(object)
    m = w[:, 0] + (w[:, 1] << 64) + (w[:, 2] << 128) + (w[:, 3] << 192)
    idx = np.arange(scalars.shape[0], dtype=object)
    return int((m * (a + idx * d)).sum()) % R * pow(O.MONT_R, -1, R) % R


def _msm_host(scalars, bases):
    acc = np.zeros((3, 4), dtype=np.uint64)
    n = scalars.shape[0]
    Z.check(Z.parallel_msm_api(0, scalars, bases, n, acc), "parallel_msm_api")
    return _affine(acc)


def test_golden_cases():
    for c in json.load(open(GOLD))["msm"]:
        got = _msm_host(_arr(c["scalars"]), _arr(c["bases"]))
        want = None if c["expected"] is None else (int(c["expected"][0], 16), int(c["expected"][1], 16))
        assert got == want, c["name"]


def test_empty_and_zero():
    acc = np.ones((3, 4), dtype=np.uint64)
    assert Z.parallel_msm_api(0, np.zeros((1, 4), np.uint64), np.zeros((1, 8), np.uint64), 0, acc) == 0
    assert not acc.any()  # the identity, as the reference leaves it
    bases = O.g1_arith_bases(50, 3, 5)
    assert _msm_host(np.zeros((50, 4), np.uint64), bases) is None


@pytest.mark.parametrize("n", [1, 2, 31, 1000, 4096])
def test_vs_oracle_with_edge_bases(n):
    """random scalars over bases that include the identity, repeated bases
    and a base next to its negation"""
    rnd = random.Random(n)
    bases = O.g1_arith_bases(n, rnd.randrange(1, R), rnd.randrange(1, R))
    if n >= 4:
        bases[1] = 0  # identity
        bases[2] = bases[3]  # repeated base
    if n >= 8:
        bases[5] = bases[6]
        y = O.fq_value(bases[6][4:])
        bases[5][4:] = O.fq_elem(O.Q_MOD - y)  # -bases[6]
    scalars = O.gen(4400 + n, n)
    if n >= 10:
        scalars[7] = 0
        scalars[8] = O.elem(R - 1)
    assert _msm_host(scalars, bases) == O.g1_msm(scalars, bases, threads=16)


@pytest.mark.parametrize("n,width", [(100, 5), (1500, 9), (3000, 10), (10000, 11), (20000, 15)])
def test_closed_form_every_window_width(n, width):
    """random scalars at a size for each window width of the plan table that
    the other tests leave out (widths 4, 8, 10, 15 and 16 come up there too),
    against the closed form and, up to 3000 points, the oracle; widths whose
    top window holds only a few of the scalars' bits, and the switch from 11
    to 15 bits between 2^13 and 2^14 points"""
    import ctypes
    cw = ctypes.c_uint32(0)
    Z.lib().zkgpu_msm_workspace_bytes(n, ctypes.byref(cw))
    assert cw.value == width
    a, d = 0x51 + width, 0x1F2E3D4C5B6A7988
    bases = O.g1_arith_bases(n, a, d)
    scalars = O.gen(4700 + width, n)
    got = _msm_host(scalars, bases)
    assert got == O.g1_mul_gen(_weighted_sum(scalars, a, d))
    if n <= 3000:
        assert got == O.g1_msm(scalars, bases)


@pytest.mark.parametrize("n", [1 << 12, 1 << 16, 1 << 20, 700001, (1 << 21) + 4097])
def test_closed_form_large(n):
    """random scalars; the odd sizes leave partial sort chunks, slices and
    segments"""
    log_n = n.bit_length()
    a, d = 0x1234567 + log_n, 0x9876543210FEDCBA
    bases = O.g1_arith_bases(n, a, d)
    scalars = O.gen(4500 + log_n, n)
    want = O.g1_mul_gen(_weighted_sum(scalars, a, d))
    assert _msm_host(scalars, bases) == want


def _skewed(kind, n, rnd):
    """scalars whose bucket counts are far from uniform: one bucket of a
    window then holds most of the n points, and its sum is split over many
    threads"""
    if kind == "equal":
        return np.tile(O.elem(rnd.randrange(R)), (n, 1))
    if kind == "zero_one":
        bits = np.array([rnd.getrandbits(1) for _ in range(n)], dtype=np.uint64)
        return np.ascontiguousarray(np.where(bits[:, None] == 1, O.elem(1), np.uint64(0)))
    if kind == "minus_one":
        return np.tile(O.elem(R - 1), (n, 1))
    if kind == "small":  # canonical values below 2^12, the rest of every window zero
        return np.ascontiguousarray(np.concatenate([O.elem(rnd.randrange(1 << 12)) for _ in range(n)]))
    raise ValueError(kind)


@pytest.mark.parametrize("kind,log_n", [("equal", 18), ("zero_one", 18), ("minus_one", 16), ("small", 16),
                                        ("equal", 20), ("zero_one", 12), ("equal", 22), ("zero_one", 22)])
def test_skewed_scalars(kind, log_n):
    """closed form over scalar sets that put most points into a few buckets
    (at 2^22 the ordinary segments are longer than a heavy bucket's)"""
    n = 1 << log_n
    rnd = random.Random(log_n * 31 + len(kind))
    a, d = rnd.randrange(1, R), rnd.randrange(1, R)
    bases = O.g1_arith_bases(n, a, d)
    scalars = _skewed(kind, n, rnd)
    assert _msm_host(scalars, bases) == O.g1_mul_gen(_weighted_sum(scalars, a, d)), kind


def test_device_entry_point(torch_cuda):
    torch = torch_cuda
    n = 5000
    bases = O.g1_arith_bases(n, 77, 1001)
    scalars = O.gen(4600, n)
    ds = torch.from_numpy(scalars.view(np.int64).reshape(-1)).to("cuda")
    db = torch.from_numpy(bases.view(np.int64).reshape(-1)).to("cuda")
    torch.cuda.synchronize()
    acc = Z.msm_device(ptr(ds), ptr(db), n)
    assert _affine(acc) == O.g1_mul_gen(_weighted_sum(scalars, 77, 1001))
    # twice in a row on the same stream, the second smaller (workspace reuse)
    acc2 = Z.msm_device(ptr(ds), ptr(db), 17)
    assert _affine(acc2) == O.g1_msm(scalars[:17], bases[:17])


def _mix64(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_repeated_calls_identical(torch_cuda):
    """The order of a bucket's entries follows the sort's atomic
    reservations and varies from call to call; the result, a group element
    returned in affine form, does not."""
    torch = torch_cuda
    n = (1 << 20) + 77
    ds = torch.empty(n * 4, dtype=torch.int64, device="cuda")
    db = torch.empty(n * 8, dtype=torch.int64, device="cuda")
    Z.fill_random(ptr(ds), n, 515)
    Z.fill_random_points(ptr(db), n, 616, 0)
    torch.cuda.synchronize()
    first = np.asarray(Z.msm_device(ptr(ds), ptr(db), n))
    for _ in range(3):
        assert (np.asarray(Z.msm_device(ptr(ds), ptr(db), n)) == first).all()
    # all scalars equal: one bucket per window, summed in segments and merged
    ds.view(n, 4)[:] = ds.view(n, 4)[0].clone()
    torch.cuda.synchronize()
    first = np.asarray(Z.msm_device(ptr(ds), ptr(db), n))
    for _ in range(3):
        assert (np.asarray(Z.msm_device(ptr(ds), ptr(db), n)) == first).all()


def test_concurrent_callers():
    """Six host threads call the host-buffer MSM at once (more than the four
    host-call contexts, so some wait for one); every result equals the
    closed form of its own inputs."""
    import threading
    cases = []
    for t in range(6):
        n = 3000 + 977 * t
        a, d = 1000 + t, 77 + 13 * t
        cases.append((O.gen(4800 + t, n), O.g1_arith_bases(n, a, d), a, d))
    got = [None] * len(cases)

    def worker(t):
        sc, bs, _, _ = cases[t]
        acc = np.zeros((3, 4), dtype=np.uint64)
        st = Z.parallel_msm_api(0, sc, bs, sc.shape[0], acc)
        got[t] = (st, _affine(acc) if st == 0 else None)

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(len(cases))]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    for t, (sc, _, a, d) in enumerate(cases):
        assert got[t][0] == 0, t
        assert got[t][1] == O.g1_mul_gen(_weighted_sum(sc, a, d)), t


def _splitmix_sum(n, seed, start=0):
    """sum of the synthetic base multipliers k_i (include/zkgpu_ntt.h:
    SplitMix64 of (seed, start + i), low bit set), as an integer"""
    total = 0
    step = 1 << 24
    with np.errstate(over="ignore"):
        for lo in range(0, n, step):
            i = np.arange(start + lo, start + min(n, lo + step), dtype=np.uint64)
            z = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + i * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            k = (z ^ (z >> np.uint64(31))) | np.uint64(1)
            total += int((k >> np.uint64(32)).sum()) << 32
            total += int((k & np.uint64(0xFFFFFFFF)).sum())
    return total


@pytest.mark.parametrize("n", [(1 << 27) + 3, 1 << 28])
def test_pieces_at_the_largest_sizes(torch_cuda, n):
    """Above 2^27 points the MSM runs as pieces (two here, the first of
    2^26 + 2 points at 2^27 + 3) whose sums the host adds; 2^28 is the largest
    length. Equal scalars v over synthetic bases k_i G: the sum is
    (v sum_i k_i mod r) G."""
    torch = torch_cuda
    seed, v = 77, 0x2468ACE013579BDF2468ACE013579BDF
    db = torch.empty(n * 8, dtype=torch.int64, device="cuda")
    Z.fill_random_points(ptr(db), n, seed, 0)
    w = torch.from_numpy(O.elem(v).view(np.int64).reshape(1, 4).copy()).to("cuda")
    ds = w.expand(n, 4).contiguous()
    torch.cuda.synchronize()
    acc = Z.msm_device(ptr(ds), ptr(db), n)
    del ds, db
    torch.cuda.empty_cache()
    assert _affine(acc) == O.g1_mul_gen(v * _splitmix_sum(n, seed) % R)


def test_host_api_pieces(torch_cuda):
    """The host-buffer entry point stages and sums pieces of at most 2^24
    points: 2^24 + 5 points run as two."""
    torch = torch_cuda
    n, seed, v = (1 << 24) + 5, 91, 0x13579BDF2468ACE0
    db = torch.empty(n * 8, dtype=torch.int64, device="cuda")
    Z.fill_random_points(ptr(db), n, seed, 0)
    torch.cuda.synchronize()
    bases = np.ascontiguousarray(db.cpu().numpy().view(np.uint64).reshape(n, 8))
    del db
    scalars = np.ascontiguousarray(np.tile(O.elem(v), (n, 1)))
    assert _msm_host(scalars, bases) == O.g1_mul_gen(v * _splitmix_sum(n, seed) % R)


def test_fill_random_points(torch_cuda):
    """Synthetic bases: point i = k_i G with k_i from SplitMix64 (the rule in
    include/zkgpu_ntt.h), checked against the oracle's k G."""
    torch = torch_cuda
    n, seed, start = 300, 42, 1000
    d = torch.empty(n * 8, dtype=torch.int64, device="cuda")
    Z.fill_random_points(ptr(d), n, seed, start)
    torch.cuda.synchronize()
    pts = d.cpu().numpy().view(np.uint64).reshape(n, 8)
    M = (1 << 64) - 1
    for i in (0, 1, 150, n - 1):
        k = _mix64((seed * 0xD1B54A32D192ED03 + (start + i) * 0x9E3779B97F4A7C15) & M) | 1
        assert (O.fq_value(pts[i][:4]), O.fq_value(pts[i][4:])) == O.g1_mul_gen(k), i
    s = O.gen(4700, n)
    assert _msm_host(s, np.ascontiguousarray(pts)) == O.g1_msm(s, pts, threads=8)
