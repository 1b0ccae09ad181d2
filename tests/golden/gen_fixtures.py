"""Generate the golden fixtures for the BN256 Fr NTT from first principles.

Pure Python big-int arithmetic, no reference code and no oracle code: every
expected value comes from the definition of the transform,

    fwd        X_j = sum_i x_i w^(ij)
    inv        y_i = n^-1 sum_j X_j w^(-ij)
    coset_fwd  X_j = sum_i x_i (g w^j)^i            (g = 7)
    coset_inv  the exact inverse of coset_fwd: y_i = g^-i n^-1 sum_j X_j w^(-ij)

with w the primitive 2^k-th root of unity 7^((r-1)/2^k) (Fr's 2^28-adic root
squared down, reference src/main.rs:140-162), evaluated point by point with
Horner's rule. Elements are stored the way the drop-in boundary sees them:
Montgomery form x * 2^256 mod r as four little-endian u64 limbs
(reference cuda/bn256.h:10-13).

Outputs (all under tests/golden/):
  fr_vectors.json       Montgomery mul/add/sub vectors incl. edge values
  gen_prefix.json       first elements of the seeded input generator
  ntt_full_k{k}.npz     full outputs of all four ops for small k (brute force)
  ntt_sampled.json      sampled outputs at k = 16, 20, 24 (Horner per index)

Run: python tests/golden/gen_fixtures.py [--skip-k24]
"""
import argparse
import json
import os
import random

import numpy as np

R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
RM = (1 << 256) % R          # Montgomery radix mod r
RM_INV = pow(RM, -1, R)
M64 = (1 << 64) - 1
G = 7
HERE = os.path.dirname(os.path.abspath(__file__))


def mont(x):
    return x * RM % R


def unmont(x):
    return x * RM_INV % R


def limbs(x):
    return [(x >> (64 * i)) & M64 for i in range(4)]


def hexl(x):
    return [f"{v:016x}" for v in limbs(x)]


def root(k):
    w = pow(G, (R - 1) >> 28, R)
    for _ in range(k, 28):
        w = w * w % R
    return w


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def gen_value(seed, i):
    """Canonical value of synthetic element i (stored Montgomery-encoded)."""
    gamma = 0x9E3779B97F4A7C15
    for a in range(64):
        st = (seed * 0xD1B54A32D192ED03 + (i * 64 + a) * 4) & M64
        l = [_mix((st + (j + 1) * gamma) & M64) for j in range(4)]
        l[3] &= (1 << 62) - 1
        x = l[0] | (l[1] << 64) | (l[2] << 128) | (l[3] << 192)
        if x < R:
            return x
    return 0


def horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R
    return acc


def op_output(op, xs, k, j):
    """Canonical value of output j of op applied to canonical inputs xs."""
    n = 1 << k
    w = root(k)
    if op == "fwd":
        return horner(xs, pow(w, j, R))
    if op == "coset_fwd":
        return horner(xs, G * pow(w, j, R) % R)
    winv = pow(w, -1, R)
    ninv = pow(n, -1, R)
    y = horner(xs, pow(winv, j, R)) * ninv % R
    if op == "inv":
        return y
    if op == "coset_inv":
        return y * pow(pow(G, -1, R), j, R) % R
    raise ValueError(op)


OPS = ("fwd", "inv", "coset_fwd", "coset_inv")


def fr_vectors():
    rnd = random.Random(2024)
    edge = [0, 1, 2, R - 1, R - 2, RM, R - RM, (R - 1) // 2, (R + 1) // 2,
            # top 64-bit limb equal to r's, sum below r: exercises the full compare
            (R >> 192) << 192, ((R >> 192) << 192) - 1, (1 << 253), (1 << 253) - 1,
            # Montgomery-domain special values
            mont(1), mont(R - 1)]
    vals = edge + [rnd.randrange(R) for _ in range(48)]
    cases = []
    for a in vals:
        for b in (vals[rnd.randrange(len(vals))], vals[rnd.randrange(len(vals))]):
            cases.append({
                "a": hexl(a), "b": hexl(b),
                "mul": hexl(a * b * RM_INV % R),      # Montgomery product of the stored words
                "add": hexl((a + b) % R),
                "sub": hexl((a - b) % R),
            })
    # pairs whose sum lands just below r with the top limbs matching r's
    for _ in range(16):
        a = ((R >> 192) << 191) + rnd.randrange(1 << 100)
        b = R - 1 - a - rnd.randrange(1 << 60)
        cases.append({"a": hexl(a), "b": hexl(b), "mul": hexl(a * b * RM_INV % R),
                      "add": hexl((a + b) % R), "sub": hexl((a - b) % R)})
    consts = {
        "r": hexl(R), "R_mod_r": hexl(RM), "R2_mod_r": hexl(RM * RM % R),
        "neg_r_inv_mod_2_64": f"{(-pow(R, -1, 1 << 64)) % (1 << 64):016x}",
        "root28_mont": hexl(mont(root(28))),
        "g_mont": hexl(mont(G)), "g_inv_mont": hexl(mont(pow(G, -1, R))),
        "roots_mont": {str(k): hexl(mont(root(k))) for k in range(1, 29)},
        "root_invs_mont": {str(k): hexl(mont(pow(root(k), -1, R))) for k in range(1, 29)},
        "n_inv_mont": {str(k): hexl(mont(pow(1 << k, -1, R))) for k in range(0, 29)},
    }
    return {"consts": consts, "cases": cases}


def seed_for(k):
    return 100 + k


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-k24", action="store_true")
    ap.add_argument("--full-k", default="1,2,3,4,5,6,7,8,9,10,11,12")
    args = ap.parse_args()

    with open(os.path.join(HERE, "fr_vectors.json"), "w") as f:
        json.dump(fr_vectors(), f, indent=0)
    print("fr_vectors.json")

    prefix = {}
    for seed in (1, 7, 1001, 2001, 3001):
        prefix[str(seed)] = [hexl(mont(gen_value(seed, i))) for i in range(16)]
    with open(os.path.join(HERE, "gen_prefix.json"), "w") as f:
        json.dump(prefix, f, indent=0)
    print("gen_prefix.json")

    for k in [int(v) for v in args.full_k.split(",")]:
        n = 1 << k
        seed = seed_for(k)
        xs = [gen_value(seed, i) for i in range(n)]
        out = {"seed": np.array([seed], dtype=np.uint64)}
        for op in OPS:
            ys = [op_output(op, xs, k, j) for j in range(n)]
            out[op] = np.array([limbs(mont(y)) for y in ys], dtype=np.uint64)
        np.savez(os.path.join(HERE, f"ntt_full_k{k}.npz"), **out)
        print(f"ntt_full_k{k}.npz")

    sampled = {}
    plan = {16: 1001, 20: 2001, 24: 3001}
    for k, seed in plan.items():
        if k == 24 and args.skip_k24:
            continue
        n = 1 << k
        xs = [gen_value(seed, i) for i in range(n)]
        rnd = random.Random(k)
        nsamp = 2 if k == 24 else 6
        idx = sorted(set([0, 1, n - 1] + [rnd.randrange(n) for _ in range(nsamp)]))
        entry = {"seed": seed, "log_n": k, "indices": idx, "ops": {}}
        for op in OPS:
            entry["ops"][op] = [hexl(mont(op_output(op, xs, k, j))) for j in idx]
            print(f"k={k} {op} done", flush=True)
        sampled[str(k)] = entry
    path = os.path.join(HERE, "ntt_sampled.json")
    if args.skip_k24 and os.path.exists(path):
        with open(path) as f:
            old = json.load(f)
        if "24" in old:
            sampled["24"] = old["24"]
    with open(path, "w") as f:
        json.dump(sampled, f, indent=0)
    print("ntt_sampled.json")


if __name__ == "__main__":
    main()
