"""Writes tests/golden/g1_vectors.json: BN256 G1 known answers computed here
with plain Python integers and affine formulas (slope with a modular
inverse), independently of the C oracle and of the library:

  * multiples k G of the generator G = (1, 2) of y^2 = x^3 + 3 over Fq,
    including k = r - 1 (= -G) and k = r (the identity);
  * small multi-scalar multiplications sum_i s_i P_i, each expected value
    computed point by point, with the edge cases of the reference's curve
    code (cuda/curve.cuh:106-297): a zero scalar, the identity as a base
    (x = y = 0), the same base twice (the doubling branch), a base and its
    negation (the identity branch), scalars 1 and r - 1.

Scalars are stored as halo2 Fr Montgomery words, bases as halo2 Fq
Montgomery words (x then y), as parallel_msm_api takes them.

    python tests/golden/gen_g1_fixtures.py
"""
import json
import os
import random

Q = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
B = 3
G = (1, 2)
RQ = (1 << 256) % Q
RR = (1 << 256) % R


def on_curve(p):
    return p is None or (p[1] * p[1] - p[0] ** 3 - B) % Q == 0


def add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    if p[0] == q[0]:
        if (p[1] + q[1]) % Q == 0:
            return None
        lam = 3 * p[0] * p[0] * pow(2 * p[1], -1, Q) % Q
    else:
        lam = (q[1] - p[1]) * pow(q[0] - p[0], -1, Q) % Q
    x = (lam * lam - p[0] - q[0]) % Q
    return (x, (lam * (p[0] - x) - p[1]) % Q)


def mul(k, p):
    acc = None
    for bit in bin(k)[2:]:
        acc = add(acc, acc)
        if bit == "1":
            acc = add(acc, p)
    return acc


def neg(p):
    return None if p is None else (p[0], (-p[1]) % Q)


def words(x):
    return [f"{(x >> (64 * i)) & (2**64 - 1):016x}" for i in range(4)]


def fr_mont(s):
    return words(s * RR % R)


def base_words(p):
    if p is None:
        return words(0) + words(0)
    return words(p[0] * RQ % Q) + words(p[1] * RQ % Q)


def point(p):
    return None if p is None else [hex(p[0]), hex(p[1])]


def main():
    rnd = random.Random(20240)
    assert on_curve(G)
    mults = {}
    ks = [1, 2, 3, 7, 12345, (1 << 128) + 1, R - 1, R, R + 1] + [rnd.randrange(R) for _ in range(4)]
    for k in ks:
        p = mul(k, G)
        assert on_curve(p)
        mults[hex(k)] = point(p)
    assert mults[hex(R)] is None and mults[hex(R - 1)] == point(neg(G))

    pts = [mul(rnd.randrange(1, R), G) for _ in range(40)]
    cases = []

    def case(name, scalars, bases):
        acc = None
        for s, p in zip(scalars, bases):
            acc = add(acc, mul(s % R, p) if p is not None else None)
        cases.append({"name": name, "scalars": [fr_mont(s) for s in scalars],
                      "bases": [base_words(p) for p in bases], "expected": point(acc)})

    case("one", [rnd.randrange(R)], [pts[0]])
    case("scalar_one", [1], [pts[1]])
    case("scalar_r_minus_1", [R - 1], [pts[2]])
    case("scalar_zero", [0, 5], [pts[3], pts[4]])
    case("all_zero", [0, 0, 0], pts[5:8])
    case("identity_base", [rnd.randrange(R), rnd.randrange(R)], [None, pts[8]])
    case("same_base_twice", [7, 7], [pts[9], pts[9]])
    case("base_and_negation", [11, 11], [pts[10], neg(pts[10])])
    case("cancel_to_identity", [3, R - 3], [pts[11], pts[11]])
    for n in (2, 3, 5, 16, 33):
        case(f"random_{n}", [rnd.randrange(R) for _ in range(n)], [pts[(12 + i) % 40] for i in range(n)])
    case("small_scalars", [rnd.randrange(1 << 20) for _ in range(24)], pts[:24])
    case("top_bits", [R - 1 - rnd.randrange(1 << 16) for _ in range(12)], pts[14:26])
    out = {"generator": [hex(G[0]), hex(G[1])], "q": hex(Q), "r": hex(R), "multiples": mults, "msm": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "g1_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {len(mults)} multiples, {len(cases)} MSM cases")


if __name__ == "__main__":
    main()
