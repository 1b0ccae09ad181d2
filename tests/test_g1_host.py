"""The library's Fq field and G1 point code (zk_gpu_amd/csrc/fq29.hpp, the
code the MSM kernels run), compiled for the host and checked against Python
big-int arithmetic and affine point formulas, including the branches of the
reference's curve code (cuda/curve.cuh:106-297): doubling through addition,
P + (-P), the identity on either side."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.gen_g1_fixtures import G, add as aff_add, mul as aff_mul, neg as aff_neg

HERE = os.path.dirname(os.path.abspath(__file__))
Q = O.Q_MOD
U64x4 = ctypes.c_uint64 * 4
U64x8 = ctypes.c_uint64 * 8
U64x12 = ctypes.c_uint64 * 12


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("g1shim") / "libg1_shim.so"
    src = os.path.join(HERE, "native", "g1_host_shim.cpp")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-x", "c++", src, "-o", str(out),
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"], check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    L.shim_fq_is_zero.restype = ctypes.c_int
    return L


def fw(x):
    return U64x4(*[int(v) for v in O.fq_elem(x % Q)])


def fv(a):
    return O.fq_value(np.array(list(a), dtype=np.uint64))


def jac(p, z=1):
    """affine (x, y) -> Jacobian words with the given z (None -> identity)"""
    if p is None:
        return U64x12(*([0] * 12))
    x, y = p
    w = [O.fq_elem(x * z * z % Q), O.fq_elem(y * z ** 3 % Q), O.fq_elem(z)]
    return U64x12(*[int(v) for v in np.concatenate(w)])


def aff(j):
    a = np.array(list(j), dtype=np.uint64)
    return O.g1_to_affine(a)


def test_field_ops(shim):
    rnd = random.Random(1)
    vals = [0, 1, 2, Q - 1, Q - 2, (Q - 1) // 2] + [rnd.randrange(Q) for _ in range(500)]
    for _ in range(1500):
        a, b = rnd.choice(vals), rnd.choice(vals)
        o = U64x4()
        shim.shim_fq_mul(fw(a), fw(b), o)
        assert fv(o) == a * b % Q
        shim.shim_fq_add(fw(a), fw(b), o)
        assert fv(o) == (a + b) % Q
        shim.shim_fq_sub(fw(a), fw(b), o)
        assert fv(o) == (a - b) % Q
    assert shim.shim_fq_is_zero(fw(0)) == 1 and shim.shim_fq_is_zero(fw(1)) == 0


def test_point_ops(shim):
    rnd = random.Random(2)
    R = O.R_MOD
    for _ in range(60):
        p = aff_mul(rnd.randrange(1, R), G)
        q = aff_mul(rnd.randrange(1, R), G)
        zp, zq = rnd.randrange(1, Q), rnd.randrange(1, Q)
        o = U64x12()
        shim.shim_g1_double(jac(p, zp), o)
        assert aff(o) == aff_add(p, p)
        shim.shim_g1_add(jac(p, zp), jac(q, zq), o)
        assert aff(o) == aff_add(p, q)
        qa = U64x8(*[int(v) for v in np.concatenate([O.fq_elem(q[0]), O.fq_elem(q[1])])])
        shim.shim_g1_add_mixed(jac(p, zp), qa, 0, o)
        assert aff(o) == aff_add(p, q)
        shim.shim_g1_add_mixed(jac(p, zp), qa, 1, o)
        assert aff(o) == aff_add(p, aff_neg(q))
    # branches: P + P, P + (-P), identity on either side
    p = aff_mul(777, G)
    o = U64x12()
    shim.shim_g1_add(jac(p, 5), jac(p, 9), o)
    assert aff(o) == aff_add(p, p)
    shim.shim_g1_add(jac(p, 5), jac(aff_neg(p), 9), o)
    assert aff(o) is None
    shim.shim_g1_add(jac(None), jac(p, 3), o)
    assert aff(o) == p
    shim.shim_g1_add(jac(p, 3), jac(None), o)
    assert aff(o) == p
    pa = U64x8(*[int(v) for v in np.concatenate([O.fq_elem(p[0]), O.fq_elem(p[1])])])
    shim.shim_g1_add_mixed(jac(p, 11), pa, 0, o)
    assert aff(o) == aff_add(p, p)
    shim.shim_g1_add_mixed(jac(p, 11), pa, 1, o)
    assert aff(o) is None
    shim.shim_g1_add_mixed(jac(None), pa, 0, o)
    assert aff(o) == p
    shim.shim_g1_add_mixed(jac(p, 2), U64x8(*([0] * 8)), 0, o)  # identity base
    assert aff(o) == p
    shim.shim_g1_double(jac(None), o)
    assert aff(o) is None


def _bases_words(points):
    w = []
    for p in points:
        w += [0] * 8 if p is None else [int(v) for v in np.concatenate([O.fq_elem(p[0]), O.fq_elem(p[1])])]
    return (ctypes.c_uint64 * len(w))(*w)


def _bucket_sum(shim, points, negs):
    n = len(points)
    out = U64x12()
    trace = (ctypes.c_uint32 * (37 * n))()
    shim.shim_g1_bucket_sum(_bases_words(points), (ctypes.c_int * n)(*negs), n, out, trace)
    return aff(out), np.frombuffer(trace, dtype=np.uint32).reshape(n, 37)


def _limb_value(l):
    return sum(int(v) << (29 * i) for i, v in enumerate(l))


def test_bucket_sum(shim):
    """The MSM's bucket accumulation (g1_bucket_add): the sum of a list of
    affine bases with signs, against affine Python arithmetic, through the
    equal-x branches (a base twice in a row: doubling; a base then its
    negation: the identity, after which the sum starts again), identity bases
    and long runs. The accumulator is in XYZZ coordinates; after every step
    it holds its invariant: X, Y, ZZ, ZZZ normalized (limbs 0..7 < 2^29) with
    values below 3q (ZZ and ZZZ, multiplier outputs, below 1.07q), it stands
    for the affine sum so far (x = X / ZZ, y = Y / ZZZ), and ZZ^3 = ZZZ^2."""
    rnd = random.Random(7)
    R = O.R_MOD
    pool = [aff_mul(rnd.randrange(1, R), G) for _ in range(24)]
    cases = []
    for length in (1, 2, 3, 17, 300):
        pts = [rnd.choice(pool) for _ in range(length)]
        cases.append((pts, [rnd.randrange(2) for _ in pts]))
    p = pool[0]
    cases.append(([p, p], [0, 0]))                      # doubling
    cases.append(([p, p], [1, 1]))                      # doubling of -P
    cases.append(([p, p, pool[1]], [0, 1, 0]))          # P - P, then restart
    cases.append(([pool[2], p, p, p], [0, 0, 1, 0]))    # X + P - P + P
    cases.append(([None, p, None, pool[3]], [0, 1, 1, 0]))  # identity bases
    cases.append(([None, None], [0, 1]))
    cases.append(([p] * 40, [0] * 40))                  # 40 P through doubling then additions
    for pts, negs in cases:
        want = None
        for q, ng in zip(pts, negs):
            if q is not None:
                want = aff_add(want, aff_neg(q) if ng else q)
        got, trace = _bucket_sum(shim, pts, negs)
        assert got == want, (len(pts), negs[:6])
        rinv = pow(1 << 261, -1, Q)  # limbs hold v * 2^261 mod q
        part = None
        for q, ng, row in zip(pts, negs, trace):
            if q is not None:
                part = aff_add(part, aff_neg(q) if ng else q)
            if row[36]:
                assert part is None
                continue
            vals = []
            for c in range(4):
                l = row[9 * c: 9 * c + 9]
                assert all(int(v) < (1 << 29) for v in l[:8])
                assert _limb_value(l) < 3 * Q
                vals.append(_limb_value(l) * rinv % Q)
            X, Y, ZZ, ZZZ = vals
            assert _limb_value(row[18:27]) < 1.07 * Q and _limb_value(row[27:36]) < 1.07 * Q
            assert pow(ZZ, 3, Q) == pow(ZZZ, 2, Q)
            assert (X * pow(ZZ, -1, Q) % Q, Y * pow(ZZZ, -1, Q) % Q) == part


def test_xyzz_chain(shim):
    """The MSM's full XYZZ addition and doubling (g1_xyzz_add,
    g1_xyzz_double: the merge, running-sum and window-sum formulas) over
    random chains of operations, against affine Python arithmetic: additions
    of fresh points with random Z in either operand order, of the
    accumulator to itself (the doubling branch) and to an earlier value of
    itself (running sums T += S), doublings, and the branches P + (-P) and
    the identity on either side. After every step the accumulator holds the
    invariant the kernels rely on: limbs 0..7 below 2^29, X and Y below 3q,
    ZZ and ZZZ below 1.1q, ZZ^3 = ZZZ^2, and (X / ZZ, Y / ZZZ) the affine
    point (ZZ = 0 exactly for the identity)."""
    rnd = random.Random(11)
    R = O.R_MOD
    pool = [aff_mul(rnd.randrange(1, R), G) for _ in range(16)]
    rinv = pow(1 << 261, -1, Q)
    chains = []
    for length in (5, 40, 200):
        ops, pts = [], []
        for _ in range(length):
            op = rnd.choice([0, 0, 1, 2, 3, 4, 5, 6])
            ops.append(op)
            pts.append(rnd.choice(pool + [None]))
        chains.append((ops, pts))
    p = pool[0]
    chains.append(([0, 0], [p, aff_neg(p)]))          # P + (-P): identity
    chains.append(([0, 1], [p, p]))                   # P + P through addition
    chains.append(([0, 4, 0, 6, 0], [p, None, p, None, aff_neg(p)]))
    chains.append(([4, 0, 2, 3, 5, 0], [None, None, p, None, None, None]))
    chains.append(([0, 4, 0, 0, 5], [p, None, aff_neg(p), p, None]))  # saved + acc with acc == saved
    chains.append(([2, 3, 0], [None, None, None]))   # operations on the identity
    for ops, pts in chains:
        n = len(ops)
        words = []
        for q in pts:
            words += list(jac(q, rnd.randrange(1, Q)))
        out = U64x12()
        trace = (ctypes.c_uint32 * (37 * n))()
        shim.shim_g1_xyzz_chain((ctypes.c_uint64 * (12 * n))(*words), (ctypes.c_int * n)(*ops), n, out, trace)
        rows = np.frombuffer(trace, dtype=np.uint32).reshape(n, 37)
        acc = saved = None
        for op, q, row in zip(ops, pts, rows):
            if op == 0 or op == 1:
                acc = aff_add(acc, q)
            elif op == 2 or op == 3:
                acc = aff_add(acc, acc)
            elif op == 4:
                saved = acc
            else:
                acc = aff_add(acc, saved)
            assert bool(row[36]) == (acc is None), (ops, op)
            vals = []
            for c in range(4):
                l = row[9 * c: 9 * c + 9]
                assert all(int(v) < (1 << 29) for v in l[:8])
                assert _limb_value(l) < 3 * Q
                vals.append(_limb_value(l) * rinv % Q)
            X, Y, ZZ, ZZZ = vals
            assert _limb_value(row[18:27]) < 1.1 * Q and _limb_value(row[27:36]) < 1.1 * Q
            assert pow(ZZ, 3, Q) == pow(ZZZ, 2, Q)
            if acc is None:
                assert ZZ == 0
            else:
                assert (X * pow(ZZ, -1, Q) % Q, Y * pow(ZZZ, -1, Q) % Q) == acc
        assert aff(out) == acc


def test_base_conversion_by_shift(shim):
    """The MSM's base records: halo2's x * 2^256 mod q words to the kernels'
    x * 2^261 mod q as canonical words, by a 5-bit shift and a value
    reduction (q_from_words_shift) and by a Montgomery product with 2^266
    mod q (q_from_words): both equal 32 w mod q, for canonical words, words
    at or above q (up to 2^256 - 1) and the edge values."""
    rnd = random.Random(17)
    vals = [0, 1, Q - 1, Q, Q + 1, 2 * Q - 1, (1 << 256) - 1, (1 << 255), (1 << 254) + 12345]
    vals += [rnd.randrange(Q) for _ in range(300)] + [rnd.randrange(1 << 256) for _ in range(300)]
    U32x8 = ctypes.c_uint32 * 8
    for v in vals:
        w = U32x8(*[(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)])
        want = 32 * v % Q
        for shift in (1, 0):
            o = U32x8()
            shim.shim_fq_from_words(w, shift, o)
            assert sum(int(o[i]) << (32 * i) for i in range(8)) == want, (hex(v), shift)
