"""The library's field code (zk_gpu_amd/csrc/fr29.hpp, host_field.hpp),
compiled for the host and checked against Python big-int arithmetic and the
golden Montgomery vectors, including worst-case limb and value bounds."""
import ctypes
import json
import os
import random
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
M29 = (1 << 29) - 1
B261 = 1 << 261


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("shim") / "libfield_shim.so"
    src = os.path.join(HERE, "native", "field_host_shim.cpp")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-x", "c++", src, "-o", str(out),
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"], check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    return L


def arr32(vals, n=9):
    return (ctypes.c_uint32 * n)(*vals)


def limbs(x):
    return [(x >> (29 * i)) & M29 for i in range(8)] + [x >> 232]


def val(l):
    return sum(int(v) << (29 * i) for i, v in enumerate(l))


def words(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def call9(fn, *args):
    out = (ctypes.c_uint32 * 9)()
    fn(*args, out)
    return list(out)


def test_constants(shim):
    out = (ctypes.c_uint32 * 29)()
    shim.shim_consts(out)
    o = list(out)
    assert val(o[0:9]) == pow(2, 266, R)
    assert val(o[9:18]) == pow(2, 261, R)
    assert val(o[18:27]) == pow(2, 517, R)
    assert o[27] == (1 << 32) // ((R >> 232) + 1)
    for c in (5, 10, 20, 40):
        k = (ctypes.c_uint32 * 9)()
        shim.shim_subconst(c, k)
        k = list(k)
        assert val(k) == c * R
        assert all(v >= M29 for v in k[:8]), c


def test_mont_mul_random_and_extremes(shim):
    rnd = random.Random(1)
    cases = []
    for _ in range(3000):
        b = rnd.randrange(R)
        a_val = rnd.randrange(80 * R)
        cases.append((limbs(a_val), b))
    # redundant first operands: every limb at the documented maximum
    mx = [0xC0000000] * 9
    mx[8] = 0x30644E * 80 >> 0
    cases.append((mx, R - 1))
    cases.append(([0xC0000000] * 8 + [0], R - 1))
    cases.append(([M29] * 8 + [0x30644E * 78], R - 1))
    for al, b in cases:
        o = call9(shim.shim_mont_mul, arr32(al), arr32(limbs(b)))
        assert val(o) % R == val(al) * b * pow(B261, -1, R) % R
        assert all(v <= M29 for v in o[:8])
        assert val(o) < R * (1 + val(al) / B261) + 1


def test_sub_add_normalize_reduce(shim):
    rnd = random.Random(2)
    for c, bmax in ((5, 3), (10, 8), (20, 18), (40, 38)):
        for _ in range(500):
            y = rnd.randrange(bmax * R)
            x = rnd.randrange(bmax * R)
            xl, yl = limbs(x), limbs(y)
            d = call9(shim.shim_sub, c, arr32(xl), arr32(yl))
            assert val(d) == x + c * R - y
            assert all(v < (1 << 29) + (1 << 30) for v in d[:8])
            s = call9(shim.shim_add, arr32(xl), arr32(yl))
            assert val(s) == x + y
            n = list(s)
            a = arr32(n)
            shim.shim_normalize(a)
            assert val(list(a)) == x + y and all(v <= M29 for v in list(a)[:8])
            shim.shim_reduce_value(a)
            v = val(list(a))
            assert v % R == (x + y) % R and v < 3 * R and all(q <= M29 for q in list(a)[:8])


def test_pack_unpack_canonical(shim):
    rnd = random.Random(3)
    for _ in range(2000):
        x = rnd.randrange(1 << 256)
        w = arr32(words(x), 8)
        l = call9(shim.shim_unpack, w)
        assert val(l) == x and all(v <= M29 for v in l[:8])
        out = (ctypes.c_uint32 * 8)()
        shim.shim_pack(arr32(l), out)
        assert list(out) == words(x)
        v = rnd.randrange(80 * R)
        shim.shim_to_canonical(arr32(limbs(v)), out)
        assert sum(int(q) << (32 * i) for i, q in enumerate(out)) == v % R


def test_host_field_against_vectors(shim):
    vec = json.load(open(os.path.join(HERE, "golden", "fr_vectors.json")))
    u64 = lambda h: (ctypes.c_uint64 * 4)(*[int(x, 16) for x in h])
    for c in vec["cases"]:
        o = (ctypes.c_uint64 * 4)()
        shim.shim_host_mul(u64(c["a"]), u64(c["b"]), o)
        assert [f"{v:016x}" for v in o] == c["mul"]
    for k, h in vec["consts"]["roots_mont"].items():
        o = (ctypes.c_uint64 * 4)()
        shim.shim_root(int(k), o)
        assert [f"{v:016x}" for v in o] == h
    for k, h in vec["consts"]["root_invs_mont"].items():
        o = (ctypes.c_uint64 * 4)()
        shim.shim_host_inv(u64(vec["consts"]["roots_mont"][k]), o)
        assert [f"{v:016x}" for v in o] == h


def test_to_limbs261(shim):
    rnd = random.Random(4)
    for x in [0, 1, R - 1] + [rnd.randrange(R) for _ in range(2000)]:
        m = x * (1 << 256) % R
        w = (ctypes.c_uint64 * 4)(*[(m >> (64 * i)) & (2**64 - 1) for i in range(4)])
        out = (ctypes.c_uint32 * 9)()
        shim.shim_to_limbs261(w, out)
        assert val(list(out)) == x * B261 % R
