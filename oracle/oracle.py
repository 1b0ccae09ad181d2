"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py). ctypes view of liboracle_ntt.so, the C
restatement of halo2's best_fft over BN256 Fr (see fr_oracle.c for the
reference file:line each function follows).

Arrays are numpy uint64 of shape (n, 4): one Fr element per row, four
little-endian 64-bit limbs, Montgomery form with R = 2^256 (reference
cuda/bn256.h:10-13)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_ntt.so")
_lib = None

R_MOD = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
MONT_R = (1 << 256) % R_MOD


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        U = ctypes.c_uint64
        I = ctypes.c_int
        UI = ctypes.c_uint
        sig = {
            "fr_add": [P, P, P], "fr_sub": [P, P, P], "fr_mul": [P, P, P], "fr_inv": [P, P],
            "fr_pow": [P, P, P], "fr_to_mont": [P, P], "fr_from_mont": [P, P], "fr_root_of_unity": [P, UI],
            "oracle_serial_fft": [P, U, P, UI], "oracle_best_fft": [P, P, UI, I],
            "oracle_twiddle_lut": [P, P, UI, UI, I],
            "oracle_fft": [P, UI, P, I], "oracle_ifft": [P, UI, P, P, I],
            "oracle_coeff_to_extended_part": [P, UI, P, P, I], "oracle_coset_ifft": [P, UI, P, P, P, I],
            "oracle_gen": [P, U, U, U], "oracle_mul_many": [P, P, P, U], "oracle_add_many": [P, P, P, U],
            "oracle_sub_many": [P, P, P, U], "oracle_to_mont_many": [P, P, U], "oracle_from_mont_many": [P, P, U],
            "oracle_eval": [P, P, U, P],
            "g1_oracle_msm": [P, P, P, U, I], "g1_oracle_to_affine": [P, P], "g1_oracle_mul_gen": [P, P],
            "g1_oracle_arith_bases": [P, U, P, P], "g1_oracle_fq_mul": [P, P, P], "g1_oracle_fq_to_mont": [P, P],
            "g1_oracle_fq_from_mont": [P, P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = U if name == "oracle_twiddle_lut" else (I if name.startswith("g1_oracle_") and
                                                                  name[10:] in ("msm", "to_affine", "mul_gen")
                                                                  else None)
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def to_limbs(x):
    return np.array([(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def from_limbs(a):
    a = [int(v) for v in a]
    return a[0] | (a[1] << 64) | (a[2] << 128) | (a[3] << 192)


def elem(x):
    """canonical integer -> (1,4) Montgomery limbs"""
    return to_limbs(x * MONT_R % R_MOD).reshape(1, 4)


def value(a):
    """(4,) Montgomery limbs -> canonical integer"""
    return from_limbs(a) * pow(MONT_R, -1, R_MOD) % R_MOD


def gen(seed, n, start=0):
    out = np.empty((n, 4), dtype=np.uint64)
    lib().oracle_gen(_p(out), seed, start, n)
    return out


def root_of_unity(log_n):
    out = np.empty((1, 4), dtype=np.uint64)
    lib().fr_root_of_unity(_p(out), log_n)
    return out


def inv(a):
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(1, 4)
    out = np.empty((1, 4), dtype=np.uint64)
    lib().fr_inv(_p(out), _p(a))
    return out


def mul_many(a, b):
    out = np.empty_like(a)
    lib().oracle_mul_many(_p(out), _p(a), _p(b), a.shape[0])
    return out


def add_many(a, b):
    out = np.empty_like(a)
    lib().oracle_add_many(_p(out), _p(a), _p(b), a.shape[0])
    return out


def sub_many(a, b):
    out = np.empty_like(a)
    lib().oracle_sub_many(_p(out), _p(a), _p(b), a.shape[0])
    return out


def twiddle_lut(omega, log_n, sparse_degree=10, with_last_level=True):
    L = lib()
    n = L.oracle_twiddle_lut(None, _p(omega), log_n, sparse_degree, int(with_last_level))
    out = np.empty((n, 4), dtype=np.uint64)
    L.oracle_twiddle_lut(_p(out), _p(omega), log_n, sparse_degree, int(with_last_level))
    return out


def fft(a, log_n, omega=None, threads=1):
    a = np.ascontiguousarray(a.copy())
    omega = root_of_unity(log_n) if omega is None else omega
    lib().oracle_fft(_p(a), log_n, _p(omega), threads)
    return a


def ifft(a, log_n, omega_inv=None, divisor=None, threads=1):
    a = np.ascontiguousarray(a.copy())
    omega_inv = inv(root_of_unity(log_n)) if omega_inv is None else omega_inv
    divisor = elem(pow(1 << log_n, -1, R_MOD)) if divisor is None else divisor
    lib().oracle_ifft(_p(a), log_n, _p(omega_inv), _p(divisor), threads)
    return a


def coset_fft(a, log_n, c, omega=None, threads=1):
    a = np.ascontiguousarray(a.copy())
    omega = root_of_unity(log_n) if omega is None else omega
    lib().oracle_coeff_to_extended_part(_p(a), log_n, _p(omega), _p(c), threads)
    return a


def coset_ifft(a, log_n, c, omega_inv=None, divisor=None, threads=1):
    a = np.ascontiguousarray(a.copy())
    omega_inv = inv(root_of_unity(log_n)) if omega_inv is None else omega_inv
    divisor = elem(pow(1 << log_n, -1, R_MOD)) if divisor is None else divisor
    c_inv = inv(c)
    lib().oracle_coset_ifft(_p(a), log_n, _p(omega_inv), _p(divisor), _p(c_inv), threads)
    return a


def run_op(op, a, log_n, g=7, threads=1):
    """op in {"fwd", "inv", "coset_fwd", "coset_inv"}; g is the coset generator."""
    if op == "fwd":
        return fft(a, log_n, threads=threads)
    if op == "inv":
        return ifft(a, log_n, threads=threads)
    if op == "coset_fwd":
        return coset_fft(a, log_n, elem(g), threads=threads)
    if op == "coset_inv":
        return coset_ifft(a, log_n, elem(g), threads=threads)
    raise ValueError(op)


def evaluate(a, x):
    out = np.empty((1, 4), dtype=np.uint64)
    lib().oracle_eval(_p(out), _p(np.ascontiguousarray(a)), a.shape[0], _p(x))
    return out


# ---------------------------------------------------------------- G1 (BN256, y^2 = x^3 + 3 over Fq)

Q_MOD = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
Q_MONT_R = (1 << 256) % Q_MOD


def fq_elem(x):
    """canonical Fq integer -> (4,) Montgomery words"""
    return to_limbs(x * Q_MONT_R % Q_MOD)


def fq_value(a):
    """(4,) Montgomery words -> canonical Fq integer"""
    return from_limbs(a) * pow(Q_MONT_R, -1, Q_MOD) % Q_MOD


def g1_msm(scalars, bases, threads=1):
    """sum_i scalars[i] * bases[i]: scalars (n, 4) Fr Montgomery words, bases
    (n, 8) affine Montgomery words (x then y). Returns None for the identity,
    else the affine point as canonical integers (x, y)."""
    out = np.zeros(8, dtype=np.uint64)
    n = scalars.shape[0]
    inf = lib().g1_oracle_msm(_p(out), _p(np.ascontiguousarray(scalars)), _p(np.ascontiguousarray(bases)), n,
                              threads)
    return None if inf else (fq_value(out[:4]), fq_value(out[4:]))


def g1_to_affine(jac):
    """(12,) Jacobian Montgomery words (x, y, z) -> None or canonical (x, y)."""
    out = np.zeros(8, dtype=np.uint64)
    inf = lib().g1_oracle_to_affine(_p(out), _p(np.ascontiguousarray(jac, dtype=np.uint64).reshape(12)))
    return None if inf else (fq_value(out[:4]), fq_value(out[4:]))


def g1_mul_gen(k):
    """k G for the generator G = (1, 2): None or canonical (x, y)."""
    out = np.zeros(8, dtype=np.uint64)
    inf = lib().g1_oracle_mul_gen(_p(out), _p(to_limbs(k)))
    return None if inf else (fq_value(out[:4]), fq_value(out[4:]))


def g1_arith_bases(n, a, d):
    """(n, 8) affine Montgomery words of (a + i d) G, i < n."""
    out = np.zeros((n, 8), dtype=np.uint64)
    lib().g1_oracle_arith_bases(_p(out), n, _p(to_limbs(a)), _p(to_limbs(d)))
    return out
