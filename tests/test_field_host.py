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


def test_canonical_extremes(shim):
    """to_canonical_words over its whole contract (limbs 0..7 <= 0xC0000000,
    the multiplier's operand bound, and V < 2^261): every multiple k r of r with k r < 2^261 and its neighbours,
    the top of the range, and random unnormalized limb vectors."""
    out = (ctypes.c_uint32 * 8)()

    def canon(l):
        shim.shim_to_canonical(arr32(l), out)
        return sum(int(q) << (32 * i) for i, q in enumerate(out))

    top = 1 << 261
    for k in range(top // R + 1):
        for v in (k * R - 1, k * R, k * R + 1, k * R + R // 2):
            if 0 <= v < top:
                assert canon(limbs(v)) == v % R, (k, v)
    for v in (0, 1, top - 1, top - R, R - 1):
        assert canon(limbs(v)) == v % R
    rnd = random.Random(11)
    checked = 0
    while checked < 3000:
        l = [rnd.choice((rnd.randrange(0xC0000001), 0xC0000000, M29, 0)) for _ in range(8)]
        l.append(rnd.randrange(1 << 29))
        v = val(l)
        if v >= top:
            continue
        assert canon(l) == v % R, l
        checked += 1


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


def _sched(shim, R):
    out = (ctypes.c_uint32 * 64)()
    n = shim.shim_dif_schedule(R, out)
    o = list(out)[:n]
    nb = (n - R) // 3
    bf = [tuple(o[3 * i:3 * i + 3]) for i in range(nb)]
    return bf, o[3 * nb:]


def test_subconst_every_c(shim):
    """sub<C> constants for every C the radix-2/4/8 schedules may instantiate
    (up to MUL_IN_VMAX + 1 = 161) and beyond: K = C*r, limbs 0..7 in
    [2^29 - 1, 2^30) so that K[i] - b[i] never underflows for a normalized b
    and adds less than 2^30 (the bound the schedule assumes)."""
    shim.shim_subconst_any.restype = ctypes.c_int
    for c in range(2, 200):
        k = (ctypes.c_uint32 * 9)()
        assert shim.shim_subconst_any(c, k) == 0
        k = list(k)
        assert val(k) == c * R, c
        assert all(M29 <= v < (1 << 30) for v in k[:8]), c


def test_schedule_constants_in_range(shim):
    mc = (ctypes.c_uint32 * 2)()
    shim.shim_mul_contract(mc)
    vmax = mc[1]
    for R in (2, 4, 8):
        bf, fout = _sched(shim, R)
        assert len(bf) == (R.bit_length() - 1) * R // 2
        assert all(2 <= c <= vmax + 1 for _, _, c in bf), (R, bf)
        assert all(f in (0, 1, 2) for f in fout)


def test_sub_every_c_at_bounds(shim):
    """a + C*r - b for every C in 2..161, with the subtrahend at the top of
    its contract (normalized, V_b up to (C-1) r - 1) and the minuend's limbs
    up to 2^32 - 1 - 2^30: exact, and every limb stays below 2^32."""
    shim.shim_sub_any.restype = ctypes.c_int
    rnd = random.Random(9)
    for c in range(2, 162):
        for t in range(40):
            vb = (c - 1) * R - 1 if t == 0 else rnd.randrange((c - 1) * R)
            yl = limbs(vb)
            if t % 3 == 0:
                xl = [(1 << 32) - 1 - (1 << 30)] * 8 + [rnd.randrange(1 << 31)]
            else:
                xl = limbs(rnd.randrange(80 * R))
            o = (ctypes.c_uint32 * 9)()
            assert shim.shim_sub_any(c, arr32(xl), arr32(yl), o) == 0
            o = list(o)
            assert val(o) == val(xl) + c * R - vb, (c, t)
            assert all(o[i] <= xl[i] + (1 << 30) for i in range(8)), (c, t)


def _root(logR):
    w28 = pow(7, (R - 1) >> 28, R)
    return pow(w28, 1 << (28 - logR), R)


def _bitrev(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("radix", [2, 4, 8])
@pytest.mark.parametrize("chain", [0, 1])
def test_dft_regs_worst_case_bounds(shim, radix, chain):
    """The pass kernels' register DFT (natural order in, bit-reversed out),
    built for the host, at the edge of its input contract (normalized limbs,
    V < 6r): all inputs 6r - 1, limbs 0..7 all 2^29 - 1, alternating extremes,
    and random values below 6r. Every output is congruent to the DFT and
    meets the multiplier's operand contract (limbs <= 0xC0000000, V <= 160 r),
    which the compile-time schedule promises."""
    shim.shim_dft.restype = ctypes.c_int
    logR = radix.bit_length() - 1
    w = _root(logR)
    tab = []
    for e in range(radix):
        tab += limbs(pow(w, e, R) * B261 % R) + [0, 0, 0]
    roots = arr32(tab, 12 * radix)
    top = 6 * R - 1
    maxl = [M29] * 8 + [(top - sum(M29 << (29 * i) for i in range(8))) >> 232]
    assert val(maxl) < 6 * R
    rnd = random.Random(100 + radix + chain)
    cases = [[top] * radix, [val(maxl)] * radix, [top if i % 2 else 0 for i in range(radix)],
             [0] * (radix - 1) + [top]]
    cases += [[rnd.randrange(6 * R) for _ in range(radix)] for _ in range(300)]
    for xs in cases:
        flat = []
        for x in xs:
            flat += maxl if x == val(maxl) else limbs(x)
        out = (ctypes.c_uint32 * (9 * radix))()
        assert shim.shim_dft(radix, chain, arr32(flat, 9 * radix), roots, logR, out) == 0
        out = list(out)
        for k in range(radix):
            o = out[9 * _bitrev(k, logR):9 * _bitrev(k, logR) + 9]
            want = sum(x * pow(w, i * k, R) for i, x in enumerate(xs)) % R
            assert val(o) % R == want, (radix, chain, k, xs[:2])
            assert all(v <= 0xC0000000 for v in o) and val(o) <= 160 * R, (radix, k)


def test_mont_cols_header_is_generated():
    """The device multiplier's column chains (mont_cols.hpp) match their
    generator, and the generator's columns hold every partial product of the
    schoolbook (a_i b_j) and of the reduction (m_i p_j, j >= 1) exactly once,
    in the column i + j they belong to."""
    import importlib.util
    gen = os.path.join(HERE, "..", "zk_gpu_amd", "csrc", "gen_mont_cols.py")
    spec = importlib.util.spec_from_file_location("gen_mont_cols", gen)
    G = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(G)
    with open(G.OUT) as f:
        assert f.read() == G.render(), "mont_cols.hpp is stale: run gen_mont_cols.py"
    ab, mp = [], []
    for k in range(1, 17):
        for kind, i, j in G.column_terms(k):
            assert i + j == k and 0 <= i < 9 and 0 <= j < 9
            (ab if kind == "ab" else mp).append((i, j))
            if kind == "mp":
                assert i < k and j >= 1  # m_i exists before column k; m_i p_0 follows the digit
    # column 0 (a_0 b_0) and the nine m_k p_0 terms are written outside the chains
    assert sorted(ab + [(0, 0)]) == [(i, j) for i in range(9) for j in range(9)]
    assert sorted(mp) == [(i, j) for i in range(9) for j in range(1, 9)]
    assert len(ab) + 1 + len(mp) + 9 == 162
