"""The oracle (C restatement of halo2's best_fft, oracle/fr_oracle.c) pinned
against the golden fixtures, which tests/golden/gen_fixtures.py computed from
the definition of each transform with Python big integers (no oracle code),
and against the invariants the mathematics fixes. CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
R = O.R_MOD
OPS = ("fwd", "inv", "coset_fwd", "coset_inv")


def _hex_elem(h):
    return np.array([int(w, 16) for w in h], dtype=np.uint64).reshape(1, 4)


def _vectors():
    with open(os.path.join(GOLD, "fr_vectors.json")) as f:
        return json.load(f)


def test_field_ops_match_vectors():
    v = _vectors()
    a = np.concatenate([_hex_elem(c["a"]) for c in v["cases"]])
    b = np.concatenate([_hex_elem(c["b"]) for c in v["cases"]])
    for name, fn in (("mul", O.mul_many), ("add", O.add_many), ("sub", O.sub_many)):
        want = np.concatenate([_hex_elem(c[name]) for c in v["cases"]])
        assert (fn(a, b) == want).all(), name


def test_constants_match_vectors():
    c = _vectors()["consts"]
    assert O.from_limbs(_hex_elem(c["r"])[0]) == R
    assert (O.root_of_unity(28) == _hex_elem(c["root28_mont"])).all()
    for ks, h in c["roots_mont"].items():
        k = int(ks)
        assert (O.root_of_unity(k) == _hex_elem(h)).all(), k
        assert (O.inv(O.root_of_unity(k)) == _hex_elem(c["root_invs_mont"][ks])).all(), k
    for ks, h in c["n_inv_mont"].items():
        assert (O.elem(pow(1 << int(ks), -1, R)) == _hex_elem(h)).all(), ks
    assert (O.elem(7) == _hex_elem(c["g_mont"])).all()
    assert (O.inv(O.elem(7)) == _hex_elem(c["g_inv_mont"])).all()


def test_root_of_unity_orders():
    for k in (1, 2, 10, 16, 24, 28):
        w = O.value(O.root_of_unity(k)[0])
        assert pow(w, 1 << k, R) == 1
        assert pow(w, 1 << (k - 1), R) == R - 1  # primitive: w^(n/2) = -1


def test_generator_prefix():
    with open(os.path.join(GOLD, "gen_prefix.json")) as f:
        pre = json.load(f)
    for seed, rows in pre.items():
        got = O.gen(int(seed), len(rows))
        assert (got == np.concatenate([_hex_elem(r) for r in rows])).all(), seed


@pytest.mark.parametrize("k", list(range(1, 13)))
def test_full_outputs_small_k(k):
    z = np.load(os.path.join(GOLD, f"ntt_full_k{k}.npz"))
    x = O.gen(int(z["seed"][0]), 1 << k)
    for op in OPS:
        for threads in (1, 4):
            assert (O.run_op(op, x, k, threads=threads) == z[op]).all(), (op, threads)


@pytest.mark.parametrize("k", [16, 20, 24])
def test_sampled_outputs(k):
    with open(os.path.join(GOLD, "ntt_sampled.json")) as f:
        e = json.load(f)[str(k)]
    x = O.gen(e["seed"], 1 << k)
    idx = np.array(e["indices"])
    for op in (OPS if k < 24 else ("fwd", "coset_inv")):
        got = O.run_op(op, x, k, threads=8)
        want = np.concatenate([_hex_elem(h) for h in e["ops"][op]])
        assert (got[idx] == want).all(), op


def test_fixtures_catch_a_planted_mistake():
    """A transform with the wrong root (w^2 instead of w) or without the n^-1
    scale must disagree with the fixtures: they are not vacuous."""
    k = 8
    z = np.load(os.path.join(GOLD, f"ntt_full_k{k}.npz"))
    x = O.gen(int(z["seed"][0]), 1 << k)
    w = O.root_of_unity(k)
    assert not (O.fft(x, k, omega=O.mul_many(w, w)) == z["fwd"]).all()
    assert not (O.ifft(x, k, divisor=O.elem(1)) == z["inv"]).all()
    assert not (O.coset_fft(x, k, O.elem(5)) == z["coset_fwd"]).all()


@pytest.mark.parametrize("k", [10, 13, 14])
def test_invariants(k):
    n = 1 << k
    x = O.gen(400 + k, n)
    # round trips
    assert (O.run_op("inv", O.run_op("fwd", x, k, threads=4), k, threads=4) == x).all()
    assert (O.run_op("coset_inv", O.run_op("coset_fwd", x, k, threads=4), k, threads=4) == x).all()
    # impulse -> all ones, constant -> n * delta_0
    one = O.elem(1)
    imp = np.zeros((n, 4), dtype=np.uint64)
    imp[0] = one[0]
    assert (O.fft(imp, k) == np.repeat(one, n, axis=0)).all()
    const = np.repeat(one, n, axis=0)
    want = np.zeros((n, 4), dtype=np.uint64)
    want[0] = O.elem(n % R)[0]
    assert (O.fft(const, k, threads=4) == want).all()
    # linearity
    y = O.gen(500 + k, n)
    assert (O.fft(O.add_many(x, y), k, threads=4) == O.add_many(O.fft(x, k, threads=4), O.fft(y, k, threads=4))).all()
    # evaluation at a point: X_j = a(w^j)
    w = O.value(O.root_of_unity(k)[0])
    X = O.fft(x, k, threads=4)
    for j in (0, 1, n // 3, n - 1):
        assert (X[j] == O.evaluate(x, O.elem(pow(w, j, R)))[0]).all()


def test_convolution_theorem():
    """fft(a) * fft(b) = fft(a * b mod x^n - 1) against a schoolbook cyclic
    product in Python big integers."""
    k = 6
    n = 1 << k
    a = O.gen(11, n)
    b = O.gen(12, n)
    av = [O.value(r) for r in a]
    bv = [O.value(r) for r in b]
    c = [0] * n
    for i in range(n):
        for j in range(n):
            c[(i + j) % n] = (c[(i + j) % n] + av[i] * bv[j]) % R
    cm = np.concatenate([O.elem(v) for v in c])
    assert (O.mul_many(O.fft(a, k), O.fft(b, k)) == O.fft(cm, k)).all()


def test_twiddle_lut_format():
    """Sparse twiddle table T[i] = w^i (i < 1024), T[1024 + j] = w^(1024 j)
    (reference src/main.rs:128-138, consumed at cuda/fft.cu:72-79)."""
    k = 14
    w = O.root_of_unity(k)
    t = O.twiddle_lut(w, k)
    wv = O.value(w[0])
    assert t.shape[0] == 1024 + (1 << (k - 10))
    for i in (0, 1, 2, 1023):
        assert O.value(t[i]) == pow(wv, i, R)
    for j in (0, 1, (1 << (k - 10)) - 1):
        assert O.value(t[1024 + j]) == pow(wv, 1024 * j, R)
