/*
 * ORACLE — test infrastructure only. Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it, and only as the checker.
 *
 * A CPU restatement of the reference's CPU path for the BN256 Fr NTT:
 * halo2's best_fft / serial_fft / parallel_fft (split four-step with a
 * sparse twiddle lookup table), as mirrored by the reference's CUDA code,
 * plus the three FFI operations built on it.
 *
 *   field element   : reference cuda/bn256.h:10-13 (BigNum, 4 x u64 LE),
 *                     constants cuda/bn256.h:52-56 (FR_R, FR_INV, FR_MODULUS, R2)
 *   Montgomery mul  : reference cuda/gpu_math.h:270-332 (CIOS semantics, but
 *                     with a full-width final compare so outputs are canonical;
 *                     the reference's strict '>' at gpu_math.h:326 is not kept)
 *   add / sub       : reference cuda/gpu_math.h:239-264 (full compare, not the
 *                     top-limb shortcut of gpu_math.h:259)
 *   pow_vartime     : reference cuda/gpu_math.h:405-417
 *   bit_reverse     : reference cuda/gpu_math.h:395-403
 *   serial_fft      : reference cuda/fft.cu:11-42 (halo2 serial_fft)
 *   split_radix_fft : reference cuda/fft.cu:61-113 (halo2 split_radix_fft)
 *   shuffle/unshuffle: reference cuda/fft.cu:129-156
 *   distribute_powers: reference cuda/fft.cu:223-232
 *   op ordering     : reference cuda/fft.cu:275-405 (coset scale before the
 *                     transform, divisor after it)
 *   twiddle LUT     : format consumed at cuda/fft.cu:72-79,97-106, built by the
 *                     caller at src/main.rs:128-138 (generate_twiddle_lookup_table
 *                     with sparse degree 10, last level included)
 *   omega derivation: src/main.rs:140-162 (omega_k = ROOT_OF_UNITY^(2^(28-k)))
 *
 * halo2 itself (halo2_proofs fork + pairing/halo2curves BN256 Fr) is a
 * third-party crate absent from the reference tree, version unpinned
 * (Cargo.lock is not committed; Cargo.toml names ff 0.12). Its published
 * algorithm is restated here; parity is pinned by the golden fixtures under
 * tests/golden/ (brute-force DFT and Horner evaluations in Python big-int
 * arithmetic) and by the invariant tests.
 *
 * Threads follow halo2's scheme: log_split = floor(log2(threads)).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef struct { uint64_t l[4]; } fr;
typedef unsigned __int128 u128;

static const fr FR_MOD = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
static const fr FR_ONE = {{0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL}}; /* R mod r */
static const fr FR_R2  = {{0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}}; /* R^2 mod r */
static const uint64_t FR_NINV = 0xc2e1f593efffffffULL;                                                                    /* -r^-1 mod 2^64 */
static const fr FR_ROOT28 = {{0x9632c7c5b639feb8ULL, 0x985ce3400d0ff299ULL, 0xb2dd880001b0ecd8ULL, 0x1d69070d6d98ce29ULL}}; /* 7^((r-1)/2^28), Montgomery */
static const fr FR_ZERO = {{0, 0, 0, 0}};

static int geq(const fr *a, const fr *b) {
    for (int i = 3; i >= 0; --i) {
        if (a->l[i] > b->l[i]) return 1;
        if (a->l[i] < b->l[i]) return 0;
    }
    return 1;
}
static void sub_raw(fr *r, const fr *a, const fr *b) {
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a->l[i] - b->l[i] - borrow;
        r->l[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
}
static uint64_t add_raw(fr *r, const fr *a, const fr *b) {
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a->l[i] + b->l[i] + carry;
        r->l[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    return carry;
}

/* a + b mod r (reference gpu_math.h:255-264, with a full compare) */
void fr_add(fr *r, const fr *a, const fr *b) {
    fr s;
    uint64_t c = add_raw(&s, a, b);
    if (c || geq(&s, &FR_MOD)) sub_raw(&s, &s, &FR_MOD);
    *r = s;
}
/* a - b mod r (reference gpu_math.h:239-253) */
void fr_sub(fr *r, const fr *a, const fr *b) {
    fr d;
    if (geq(a, b)) { sub_raw(&d, a, b); }
    else { fr t; add_raw(&t, a, &FR_MOD); sub_raw(&d, &t, b); }
    *r = d;
}
/* Montgomery product a*b*2^-256 mod r, CIOS over 4 x 64-bit words
 * (reference gpu_math.h:270-332 computes the same quantity). */
void fr_mul(fr *r, const fr *a, const fr *b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        uint64_t c = 0;
        for (int j = 0; j < 4; ++j) {
            u128 p = (u128)a->l[i] * b->l[j] + t[j] + c;
            t[j] = (uint64_t)p; c = (uint64_t)(p >> 64);
        }
        u128 s = (u128)t[4] + c; t[4] = (uint64_t)s; t[5] = (uint64_t)(s >> 64);
        uint64_t m = t[0] * FR_NINV;
        u128 p = (u128)m * FR_MOD.l[0] + t[0];
        c = (uint64_t)(p >> 64);
        for (int j = 1; j < 4; ++j) {
            p = (u128)m * FR_MOD.l[j] + t[j] + c;
            t[j - 1] = (uint64_t)p; c = (uint64_t)(p >> 64);
        }
        s = (u128)t[4] + c; t[3] = (uint64_t)s; t[4] = t[5] + (uint64_t)(s >> 64);
    }
    fr out = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || geq(&out, &FR_MOD)) sub_raw(&out, &out, &FR_MOD);
    *r = out;
}
/* square-and-multiply over a 4-limb exponent (reference gpu_math.h:405-417) */
void fr_pow(fr *r, const fr *base, const uint64_t e[4]) {
    fr acc = FR_ONE;
    for (int i = 3; i >= 0; --i)
        for (int b = 63; b >= 0; --b) {
            fr_mul(&acc, &acc, &acc);
            if ((e[i] >> b) & 1) fr_mul(&acc, &acc, base);
        }
    *r = acc;
}
static void fr_pow_u64(fr *r, const fr *base, uint64_t e) {
    uint64_t ee[4] = {e, 0, 0, 0};
    fr_pow(r, base, ee);
}
/* inverse by Fermat: x^(r-2) */
void fr_inv(fr *r, const fr *a) {
    uint64_t e[4] = {FR_MOD.l[0] - 2, FR_MOD.l[1], FR_MOD.l[2], FR_MOD.l[3]};
    fr_pow(r, a, e);
}
void fr_to_mont(fr *r, const fr *a) { fr_mul(r, a, &FR_R2); }
void fr_from_mont(fr *r, const fr *a) { fr one = {{1, 0, 0, 0}}; fr_mul(r, a, &one); }

/* reference gpu_math.h:395-403 */
static uint64_t bitreverse(uint64_t n, unsigned l) {
    uint64_t r = 0;
    for (unsigned i = 0; i < l; ++i) { r = (r << 1) | (n & 1); n >>= 1; }
    return r;
}

/* omega_k = ROOT_OF_UNITY^(2^(28-k)) (reference src/main.rs:140-162) */
void fr_root_of_unity(fr *r, unsigned log_n) {
    fr w = FR_ROOT28;
    for (unsigned i = log_n; i < 28; ++i) fr_mul(&w, &w, &w);
    *r = w;
}

/* halo2 serial_fft (reference cuda/fft.cu:11-42): bit-reversal permutation
 * followed by radix-2 DIT; stage root w_m = omega^(n/2m), running twiddle. */
void oracle_serial_fft(fr *a, uint64_t n, const fr *omega, unsigned log_n) {
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t rk = bitreverse(k, log_n);
        if (k < rk) { fr t = a[k]; a[k] = a[rk]; a[rk] = t; }
    }
    uint64_t m = 1;
    for (unsigned s = 0; s < log_n; ++s) {
        fr w_m;
        fr_pow_u64(&w_m, omega, n / (2 * m));
        for (uint64_t k = 0; k < n; k += 2 * m) {
            fr w = FR_ONE;
            for (uint64_t j = 0; j < m; ++j) {
                fr t = a[k + j + m];
                fr_mul(&t, &t, &w);
                fr_sub(&a[k + j + m], &a[k + j], &t);
                fr_add(&a[k + j], &a[k + j], &t);
                fr_mul(&w, &w, &w_m);
            }
        }
        m *= 2;
    }
}

/* generate_twiddle_lookup_table(omega, log_n, sparse_degree, with_last_level)
 * as consumed at reference cuda/fft.cu:72-79,97-106: dense table omega^i of
 * length n when sparse_degree > log_n, otherwise omega^i for
 * i < 2^sparse_degree followed by omega^(2^sparse_degree * j) for
 * j < 2^(log_n - sparse_degree - !with_last_level). Returns the length. */
uint64_t oracle_twiddle_lut(fr *out, const fr *omega, unsigned log_n, unsigned sparse_degree, int with_last_level) {
    if (sparse_degree > log_n) {
        uint64_t len = 1ULL << log_n;
        fr w = FR_ONE;
        for (uint64_t i = 0; i < len; ++i) { if (out) out[i] = w; fr_mul(&w, &w, omega); }
        return len;
    }
    uint64_t low = 1ULL << sparse_degree;
    uint64_t high = 1ULL << (log_n - sparse_degree - (with_last_level ? 0 : 1));
    if (out) {
        fr w = FR_ONE;
        for (uint64_t i = 0; i < low; ++i) { out[i] = w; fr_mul(&w, &w, omega); }
        fr hw;
        fr_pow_u64(&hw, omega, low);
        w = FR_ONE;
        for (uint64_t j = 0; j < high; ++j) { out[low + j] = w; fr_mul(&w, &w, &hw); }
    }
    return low + high;
}

#define SPARSE_DEGREE 10
static void lut_lookup(fr *r, const fr *lut, uint64_t idx) {
    uint64_t lo = idx & ((1ULL << SPARSE_DEGREE) - 1), hi = idx >> SPARSE_DEGREE;
    *r = lut[lo];
    if (hi > 0) fr_mul(r, r, &lut[(1ULL << SPARSE_DEGREE) + hi]);
}

/* halo2 split_radix_fft (reference cuda/fft.cu:61-113): column sub_fft_offset
 * of the split, gathered in bit-reversed order, a split_m-point DIT whose
 * stage roots come from the sparse LUT, then output t scaled by
 * omega^(sub_fft_offset * t). */
static void split_radix_fft(fr *tmp, const fr *a, const fr *lut, uint64_t n, uint64_t sub_fft_offset, unsigned log_split) {
    uint64_t split_m = 1ULL << log_split, sub_n = n >> log_split;
    for (uint64_t i = 0; i < split_m; ++i) tmp[bitreverse(i, log_split)] = a[i * sub_n + sub_fft_offset];
    uint64_t m = 1;
    for (unsigned s = 0; s < log_split; ++s) {
        fr w_m;
        lut_lookup(&w_m, lut, sub_n * split_m / (2 * m));
        for (uint64_t k = 0; k < split_m; k += 2 * m) {
            fr w = FR_ONE;
            for (uint64_t j = 0; j < m; ++j) {
                fr t = tmp[k + j + m];
                fr_mul(&t, &t, &w);
                fr_sub(&tmp[k + j + m], &tmp[k + j], &t);
                fr_add(&tmp[k + j], &tmp[k + j], &t);
                fr_mul(&w, &w, &w_m);
            }
        }
        m *= 2;
    }
    fr omega;
    lut_lookup(&omega, lut, sub_fft_offset);
    fr w = FR_ONE;
    for (uint64_t i = 0; i < split_m; ++i) { fr_mul(&tmp[i], &tmp[i], &w); fr_mul(&w, &w, &omega); }
}

typedef struct {
    int phase;
    fr *a, *tmp;
    const fr *lut;
    fr omega;
    uint64_t n, lo, hi;
    unsigned log_n, log_split;
} job_t;

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    uint64_t split_m = 1ULL << j->log_split, sub_n = j->n >> j->log_split;
    if (j->phase == 0) {
        /* split fft over columns [lo, hi) */
        for (uint64_t c = j->lo; c < j->hi; ++c) split_radix_fft(j->tmp + c * split_m, j->a, j->lut, j->n, c, j->log_split);
    } else if (j->phase == 1) {
        /* shuffle (reference cuda/fft.cu:129-135) */
        for (uint64_t idx = j->lo; idx < j->hi; ++idx) j->a[idx] = j->tmp[(idx % sub_n) * split_m + idx / sub_n];
    } else if (j->phase == 2) {
        /* sub fft rows [lo, hi) with omega^split_m */
        for (uint64_t r = j->lo; r < j->hi; ++r) oracle_serial_fft(j->a + r * sub_n, sub_n, &j->omega, j->log_n - j->log_split);
    } else if (j->phase == 3) {
        /* copy (reference cuda/fft.cu:143-147) */
        for (uint64_t idx = j->lo; idx < j->hi; ++idx) j->tmp[idx] = j->a[idx];
    } else {
        /* unshuffle (reference cuda/fft.cu:149-156) */
        uint64_t mask = split_m - 1;
        for (uint64_t idx = j->lo; idx < j->hi; ++idx) j->a[idx] = j->tmp[sub_n * (idx & mask) + (idx >> j->log_split)];
    }
    return NULL;
}

static void run_phase(job_t *proto, int phase, uint64_t total, int threads) {
    pthread_t th[256];
    job_t jobs[256];
    if (threads > 256) threads = 256;
    uint64_t chunk = (total + threads - 1) / threads;
    int started = 0;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = *proto;
        jobs[t].phase = phase;
        jobs[t].lo = t * chunk;
        jobs[t].hi = (t + 1) * chunk < total ? (t + 1) * chunk : total;
        if (jobs[t].lo >= jobs[t].hi) break;
        if (threads == 1) { worker(&jobs[t]); continue; }
        pthread_create(&th[t], NULL, worker, &jobs[t]);
        started = t + 1;
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
}

/* halo2 parallel_fft: split four-step with log_split = floor(log2 threads) */
static void parallel_fft(fr *a, const fr *omega, unsigned log_n, unsigned log_split, int threads) {
    uint64_t n = 1ULL << log_n, split_m = 1ULL << log_split, sub_n = n >> log_split;
    uint64_t lut_len = oracle_twiddle_lut(NULL, omega, log_n, SPARSE_DEGREE, 1);
    fr *lut = (fr *)malloc(lut_len * sizeof(fr));
    oracle_twiddle_lut(lut, omega, log_n, SPARSE_DEGREE, 1);
    fr *tmp = (fr *)malloc(n * sizeof(fr));
    job_t proto;
    memset(&proto, 0, sizeof(proto));
    proto.a = a; proto.tmp = tmp; proto.lut = lut; proto.n = n; proto.log_n = log_n; proto.log_split = log_split;
    run_phase(&proto, 0, sub_n, threads);
    run_phase(&proto, 1, n, threads);
    fr_pow_u64(&proto.omega, omega, split_m);
    run_phase(&proto, 2, split_m, threads);
    run_phase(&proto, 3, n, threads);
    run_phase(&proto, 4, n, threads);
    free(tmp);
    free(lut);
}

static unsigned log2_floor(unsigned x) { unsigned r = 0; while ((1u << (r + 1)) <= x) ++r; return r; }

/* halo2 best_fft: serial when log_n <= log2(threads), else the split
 * parallel transform. The sparse LUT needs log_n >= 10 (it indexes
 * lut[1024 + hi]); smaller transforms take the serial path. */
void oracle_best_fft(fr *a, const fr *omega, unsigned log_n, int threads) {
    if (threads < 1) threads = 1;
    unsigned log_threads = log2_floor((unsigned)threads);
    uint64_t n = 1ULL << log_n;
    if (log_n <= log_threads || log_n < SPARSE_DEGREE || log_threads == 0) oracle_serial_fft(a, n, omega, log_n);
    else parallel_fft(a, omega, log_n, log_threads, threads);
}

/* a[i] *= c^i (reference cuda/fft.cu:223-232, distribute_powers) */
static void distribute_powers(fr *a, uint64_t n, const fr *c) {
    fr w = FR_ONE;
    for (uint64_t i = 0; i < n; ++i) { fr_mul(&a[i], &a[i], &w); fr_mul(&w, &w, c); }
}

/* The three FFI operations of reference cuda/fft.cu:560-576, in the op
 * order of do_parallel_fft (cuda/fft.cu:275-405). */
void oracle_fft(fr *a, unsigned log_n, const fr *omega, int threads) { oracle_best_fft(a, omega, log_n, threads); }
void oracle_ifft(fr *a, unsigned log_n, const fr *omega_inv, const fr *divisor, int threads) {
    uint64_t n = 1ULL << log_n;
    oracle_best_fft(a, omega_inv, log_n, threads);
    for (uint64_t i = 0; i < n; ++i) fr_mul(&a[i], &a[i], divisor);
}
void oracle_coeff_to_extended_part(fr *a, unsigned log_n, const fr *omega, const fr *c, int threads) {
    distribute_powers(a, 1ULL << log_n, c);
    oracle_best_fft(a, omega, log_n, threads);
}
/* Coset inverse (no reference entry point): the inverse transform followed
 * by a[i] *= c_inv^i, the exact inverse of oracle_coeff_to_extended_part. */
void oracle_coset_ifft(fr *a, unsigned log_n, const fr *omega_inv, const fr *divisor, const fr *c_inv, int threads) {
    oracle_ifft(a, log_n, omega_inv, divisor, threads);
    distribute_powers(a, 1ULL << log_n, c_inv);
}

/* Seeded synthetic elements: element i draws attempts a = 0, 1, ...; each
 * attempt takes four SplitMix64 outputs from the state
 * seed * 0xD1B54A32D192ED03 + (i * 64 + a) * 4, masks the top limb to 62
 * bits and accepts the value if it is below r. The accepted value is
 * Montgomery-encoded. */
static uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
void oracle_gen_element(fr *out, uint64_t seed, uint64_t i) {
    const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
    for (uint64_t a = 0; a < 64; ++a) {
        uint64_t st = seed * 0xD1B54A32D192ED03ULL + (i * 64 + a) * 4;
        fr x;
        for (int j = 0; j < 4; ++j) x.l[j] = splitmix_mix(st + (uint64_t)(j + 1) * gamma);
        x.l[3] &= (1ULL << 62) - 1;
        if (!geq(&x, &FR_MOD)) { fr_to_mont(out, &x); return; }
    }
    *out = FR_ZERO;
}
void oracle_gen(fr *out, uint64_t seed, uint64_t start, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) oracle_gen_element(&out[i], seed, start + i);
}

/* batch helpers for tests */
void oracle_mul_many(fr *r, const fr *a, const fr *b, uint64_t n) { for (uint64_t i = 0; i < n; ++i) fr_mul(&r[i], &a[i], &b[i]); }
void oracle_add_many(fr *r, const fr *a, const fr *b, uint64_t n) { for (uint64_t i = 0; i < n; ++i) fr_add(&r[i], &a[i], &b[i]); }
void oracle_sub_many(fr *r, const fr *a, const fr *b, uint64_t n) { for (uint64_t i = 0; i < n; ++i) fr_sub(&r[i], &a[i], &b[i]); }
void oracle_to_mont_many(fr *r, const fr *a, uint64_t n) { for (uint64_t i = 0; i < n; ++i) fr_to_mont(&r[i], &a[i]); }
void oracle_from_mont_many(fr *r, const fr *a, uint64_t n) { for (uint64_t i = 0; i < n; ++i) fr_from_mont(&r[i], &a[i]); }
/* Horner evaluation of sum_i a[i] * x^i (all Montgomery form) */
void oracle_eval(fr *r, const fr *a, uint64_t n, const fr *x) {
    fr acc = FR_ZERO;
    for (uint64_t i = n; i-- > 0;) { fr_mul(&acc, &acc, x); fr_add(&acc, &acc, &a[i]); }
    *r = acc;
}
