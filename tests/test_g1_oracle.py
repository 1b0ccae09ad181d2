"""The G1 oracle (oracle/g1_oracle.c, the CPU restatement of the reference's
MSM) against known answers computed with Python integers and affine formulas
(tests/golden/g1_vectors.json), the group order, and the closed form of an
MSM over bases in arithmetic progression."""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "g1_vectors.json")
R = O.R_MOD
Q = O.Q_MOD


def _vec():
    return json.load(open(GOLD))


def _arr(rows):
    return np.array([[int(w, 16) for w in row] for row in rows], dtype=np.uint64)


def _point(p):
    return None if p is None else (int(p[0], 16), int(p[1], 16))


def test_generator_multiples():
    v = _vec()
    assert (int(v["q"], 16), int(v["r"], 16)) == (Q, R)
    for k, p in v["multiples"].items():
        assert O.g1_mul_gen(int(k, 16) % (1 << 256)) == _point(p), k
    assert O.g1_mul_gen(R) is None  # the group order
    x, y = O.g1_mul_gen(12345)
    assert (y * y - x ** 3 - 3) % Q == 0


@pytest.mark.parametrize("threads", [1, 3])
def test_msm_cases(threads):
    for c in _vec()["msm"]:
        got = O.g1_msm(_arr(c["scalars"]), _arr(c["bases"]), threads=threads)
        assert got == _point(c["expected"]), c["name"]


def test_arith_bases_closed_form():
    """bases (a + i d) G: sum_i s_i (a + i d) G = (sum_i s_i (a + i d) mod r) G."""
    rnd = random.Random(5)
    n = 300
    a, d = rnd.randrange(1, R), rnd.randrange(1, R)
    bases = O.g1_arith_bases(n, a, d)
    for i in (0, 1, 77, n - 1):
        x, y = O.fq_value(bases[i][:4]), O.fq_value(bases[i][4:])
        assert (x, y) == O.g1_mul_gen((a + i * d) % R)
    s = [rnd.randrange(R) for _ in range(n)]
    scal = np.concatenate([O.elem(v) for v in s])
    want = O.g1_mul_gen(sum(si * (a + i * d) for i, si in enumerate(s)) % R)
    assert O.g1_msm(scal, bases, threads=4) == want


def test_planted_mistake_fails():
    """A wrong scalar (one bit flipped) or a wrong base changes the result."""
    c = [c for c in _vec()["msm"] if c["name"] == "random_16"][0]
    sc, bs = _arr(c["scalars"]), _arr(c["bases"])
    want = _point(c["expected"])
    assert O.g1_msm(sc, bs) == want
    sc2 = sc.copy()
    sc2[3, 0] ^= np.uint64(1 << 5)
    assert O.g1_msm(sc2, bs) != want
    bs2 = bs.copy()
    bs2[[4, 5]] = bs2[[5, 4]]
    assert O.g1_msm(sc, bs2) != want


def test_to_affine_of_jacobian():
    """(x z^2, y z^3, z) maps back to (x, y); z = 0 is the identity."""
    x, y = O.g1_mul_gen(99)
    z = 123456789
    jac = np.concatenate([O.fq_elem(x * z * z % Q), O.fq_elem(y * z ** 3 % Q), O.fq_elem(z)])
    assert O.g1_to_affine(jac) == (x, y)
    assert O.g1_to_affine(np.zeros(12, dtype=np.uint64)) is None
