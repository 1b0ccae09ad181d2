/*
 * ORACLE — test infrastructure only. Nothing in the product library links,
 * loads or calls this file; only tests/ and the cpu_baseline leg of bench.py
 * may use it, and only as the checker.
 *
 * CPU restatement of the reference's G1 multi-scalar multiplication
 * (parallel_msm_api, cuda/fft.cu:450-551 and 578-582; curve arithmetic
 * cuda/curve.cuh:13-297; field constants cuda/bn256.h:35-45):
 *
 *   result = sum_i scalar_i * base_i over BN256 G1, y^2 = x^3 + 3 over Fq,
 *   scalars read as canonical integers of their Fr Montgomery words
 *   (get_at's fr_mul(coeff, ONE), curve.cuh:75-99), affine bases with
 *   x = y = 0 as the identity (x_y_is_zero, curve.cuh:62-69).
 *
 * The reference accumulates per-thread buckets of c-bit windows and sums the
 * threads' results in a tree; the group element is the same for any
 * summation order. This restatement uses a different order on purpose:
 * bit-serial Straus (one doubling per scalar bit, one mixed addition per set
 * bit), split over threads by point ranges. Jacobian formulas dbl-2009-l,
 * madd-2007-bl and add-2007-bl, as in curve.cuh:106-297. Fq: 4 x 64-bit
 * Montgomery CIOS (R = 2^256), FQ_INV = -q^-1 mod 2^64 (bn256.h:38).
 *
 * Pinned by: tests/golden/g1_vectors.json (Python big-int affine
 * arithmetic: multiples of the generator (1, 2) and small MSMs), the group
 * order (r G = O), and the closed form sum_i s_i (a + i d) G = (sum_i s_i (a +
 * i d) mod r) G on generated bases.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t l[4]; } fq_t;
typedef unsigned __int128 u128;

static const fq_t Q_MOD = {{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
static const uint64_t Q_INV = 0x87d20782e4866389ULL;                                                                 /* -q^-1 mod 2^64 */
static const fq_t Q_ONE = {{0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL}}; /* 2^256 mod q */
static const fq_t Q_R2 = {{0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL}};  /* 2^512 mod q */
/* Fr for scalar conversion */
static const fq_t R_MOD = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
static const uint64_t R_INV = 0xc2e1f593efffffffULL;

static int geq4(const fq_t *a, const fq_t *b) {
    for (int i = 3; i >= 0; --i) {
        if (a->l[i] > b->l[i]) return 1;
        if (a->l[i] < b->l[i]) return 0;
    }
    return 1;
}
static void sub4(fq_t *r, const fq_t *a, const fq_t *b) {
    uint64_t bw = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a->l[i] - b->l[i] - bw;
        r->l[i] = (uint64_t)d;
        bw = (uint64_t)(d >> 64) & 1;
    }
}
static uint64_t add4(fq_t *r, const fq_t *a, const fq_t *b) {
    uint64_t c = 0;
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a->l[i] + b->l[i] + c;
        r->l[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return c;
}
/* Montgomery product a * b * 2^-256 mod m, canonical output (CIOS) */
static void mont(fq_t *o, const fq_t *a, const fq_t *b, const fq_t *m, uint64_t minv) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        uint64_t c = 0;
        for (int j = 0; j < 4; ++j) {
            u128 p = (u128)a->l[i] * b->l[j] + t[j] + c;
            t[j] = (uint64_t)p;
            c = (uint64_t)(p >> 64);
        }
        u128 s = (u128)t[4] + c;
        t[4] = (uint64_t)s;
        t[5] = (uint64_t)(s >> 64);
        uint64_t q = t[0] * minv;
        u128 p = (u128)q * m->l[0] + t[0];
        c = (uint64_t)(p >> 64);
        for (int j = 1; j < 4; ++j) {
            p = (u128)q * m->l[j] + t[j] + c;
            t[j - 1] = (uint64_t)p;
            c = (uint64_t)(p >> 64);
        }
        s = (u128)t[4] + c;
        t[3] = (uint64_t)s;
        t[4] = t[5] + (uint64_t)(s >> 64);
    }
    fq_t r = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || geq4(&r, m)) sub4(&r, &r, m);
    *o = r;
}
static void fmul(fq_t *o, const fq_t *a, const fq_t *b) { mont(o, a, b, &Q_MOD, Q_INV); }
static void fadd(fq_t *o, const fq_t *a, const fq_t *b) {
    fq_t r;
    uint64_t c = add4(&r, a, b);
    if (c || geq4(&r, &Q_MOD)) sub4(&r, &r, &Q_MOD);
    *o = r;
}
static void fsub(fq_t *o, const fq_t *a, const fq_t *b) {
    fq_t r;
    if (geq4(a, b)) {
        sub4(&r, a, b);
    } else {
        fq_t t;
        add4(&t, a, &Q_MOD);
        sub4(&r, &t, b);
    }
    *o = r;
}
static int fzero(const fq_t *a) { return (a->l[0] | a->l[1] | a->l[2] | a->l[3]) == 0; }
static void finv(fq_t *o, const fq_t *a) {
    fq_t e = Q_MOD, acc = Q_ONE;
    e.l[0] -= 2;
    for (int i = 3; i >= 0; --i)
        for (int k = 63; k >= 0; --k) {
            fmul(&acc, &acc, &acc);
            if ((e.l[i] >> k) & 1) fmul(&acc, &acc, a);
        }
    *o = acc;
}

typedef struct { fq_t x, y, z; } jac_t;
static const jac_t J_INF = {{{0, 0, 0, 0}}, {{0, 0, 0, 0}}, {{0, 0, 0, 0}}};

static void jdbl(jac_t *o, const jac_t *p) {
    if (fzero(&p->z)) { *o = *p; return; }
    fq_t A, B, C, D, E, F, t;
    fmul(&A, &p->x, &p->x);
    fmul(&B, &p->y, &p->y);
    fmul(&C, &B, &B);
    fadd(&t, &p->x, &B);
    fmul(&t, &t, &t);
    fsub(&t, &t, &A);
    fsub(&t, &t, &C);
    fadd(&D, &t, &t);
    fadd(&E, &A, &A);
    fadd(&E, &E, &A);
    fmul(&F, &E, &E);
    jac_t r;
    fadd(&t, &D, &D);
    fsub(&r.x, &F, &t);
    fq_t c8;
    fadd(&c8, &C, &C);
    fadd(&c8, &c8, &c8);
    fadd(&c8, &c8, &c8);
    fsub(&t, &D, &r.x);
    fmul(&t, &E, &t);
    fsub(&r.y, &t, &c8);
    fmul(&t, &p->y, &p->z);
    fadd(&r.z, &t, &t);
    *o = r;
}
/* p + (x2, y2) affine */
static void jmadd(jac_t *o, const jac_t *p, const fq_t *x2, const fq_t *y2) {
    if (fzero(&p->z)) { o->x = *x2; o->y = *y2; o->z = Q_ONE; return; }
    fq_t Z1Z1, U2, S2, H, HH, I, J, R, V, t;
    fmul(&Z1Z1, &p->z, &p->z);
    fmul(&U2, x2, &Z1Z1);
    fmul(&t, &p->z, &Z1Z1);
    fmul(&S2, y2, &t);
    fsub(&H, &U2, &p->x);
    fsub(&R, &S2, &p->y);
    fadd(&R, &R, &R);
    if (fzero(&H)) {
        if (fzero(&R)) { jdbl(o, p); return; }
        *o = J_INF;
        return;
    }
    fmul(&HH, &H, &H);
    fadd(&I, &HH, &HH);
    fadd(&I, &I, &I);
    fmul(&J, &H, &I);
    fmul(&V, &p->x, &I);
    jac_t r;
    fmul(&t, &R, &R);
    fsub(&t, &t, &J);
    fsub(&t, &t, &V);
    fsub(&r.x, &t, &V);
    fsub(&t, &V, &r.x);
    fmul(&t, &R, &t);
    fq_t yj;
    fmul(&yj, &p->y, &J);
    fadd(&yj, &yj, &yj);
    fsub(&r.y, &t, &yj);
    fadd(&t, &p->z, &H);
    fmul(&t, &t, &t);
    fsub(&t, &t, &Z1Z1);
    fsub(&r.z, &t, &HH);
    *o = r;
}
static void jadd(jac_t *o, const jac_t *p, const jac_t *q) {
    if (fzero(&p->z)) { *o = *q; return; }
    if (fzero(&q->z)) { *o = *p; return; }
    fq_t Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, R, V, t;
    fmul(&Z1Z1, &p->z, &p->z);
    fmul(&Z2Z2, &q->z, &q->z);
    fmul(&U1, &p->x, &Z2Z2);
    fmul(&U2, &q->x, &Z1Z1);
    fmul(&t, &q->z, &Z2Z2);
    fmul(&S1, &p->y, &t);
    fmul(&t, &p->z, &Z1Z1);
    fmul(&S2, &q->y, &t);
    fsub(&H, &U2, &U1);
    fsub(&R, &S2, &S1);
    fadd(&R, &R, &R);
    if (fzero(&H)) {
        if (fzero(&R)) { jdbl(o, p); return; }
        *o = J_INF;
        return;
    }
    fadd(&t, &H, &H);
    fmul(&I, &t, &t);
    fmul(&J, &H, &I);
    fmul(&V, &U1, &I);
    jac_t r;
    fmul(&t, &R, &R);
    fsub(&t, &t, &J);
    fsub(&t, &t, &V);
    fsub(&r.x, &t, &V);
    fsub(&t, &V, &r.x);
    fmul(&t, &R, &t);
    fq_t sj;
    fmul(&sj, &S1, &J);
    fadd(&sj, &sj, &sj);
    fsub(&r.y, &t, &sj);
    fadd(&t, &p->z, &q->z);
    fmul(&t, &t, &t);
    fsub(&t, &t, &Z1Z1);
    fsub(&t, &t, &Z2Z2);
    fmul(&r.z, &t, &H);
    *o = r;
}
/* affine Montgomery words (x, y), inf flag */
static void to_affine(const jac_t *p, fq_t *x, fq_t *y, int *inf) {
    if (fzero(&p->z)) { memset(x, 0, 32); memset(y, 0, 32); *inf = 1; return; }
    fq_t zi, zi2, zi3;
    finv(&zi, &p->z);
    fmul(&zi2, &zi, &zi);
    fmul(&zi3, &zi2, &zi);
    fmul(x, &p->x, &zi2);
    fmul(y, &p->y, &zi3);
    *inf = 0;
}
/* canonical integer of a Montgomery Fr element (get_at's fr_mul(coeff, ONE)) */
static void fr_canonical(fq_t *o, const uint64_t *w) {
    fq_t a, one = {{1, 0, 0, 0}};
    memcpy(a.l, w, 32);
    mont(o, &a, &one, &R_MOD, R_INV);
}

/* ---------------------------------------------------------------- MSM */

typedef struct {
    const uint64_t *scalars, *bases;
    uint64_t lo, hi;
    jac_t acc;
} msm_job;

static void *msm_worker(void *arg) {
    msm_job *j = (msm_job *)arg;
    const uint64_t n = j->hi - j->lo;
    fq_t *s = (fq_t *)malloc(sizeof(fq_t) * (n ? n : 1));
    for (uint64_t i = 0; i < n; ++i) fr_canonical(&s[i], j->scalars + 4 * (j->lo + i));
    jac_t acc = J_INF;
    for (int bit = 255; bit >= 0; --bit) {
        jdbl(&acc, &acc);
        for (uint64_t i = 0; i < n; ++i) {
            if (!((s[i].l[bit >> 6] >> (bit & 63)) & 1)) continue;
            const uint64_t *b = j->bases + 8 * (j->lo + i);
            fq_t x, y;
            memcpy(x.l, b, 32);
            memcpy(y.l, b + 4, 32);
            if (fzero(&x) && fzero(&y)) continue; /* identity base */
            jmadd(&acc, &acc, &x, &y);
        }
    }
    j->acc = acc;
    free(s);
    return NULL;
}

/* out = x, y (Montgomery words, zero for the identity); returns 1 for the identity */
int g1_oracle_msm(uint64_t *out, const uint64_t *scalars, const uint64_t *bases, uint64_t n, int threads) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n) threads = n ? (int)n : 1;
    msm_job *jobs = (msm_job *)calloc((size_t)threads, sizeof(msm_job));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        jobs[t].scalars = scalars;
        jobs[t].bases = bases;
        jobs[t].lo = n * (uint64_t)t / (uint64_t)threads;
        jobs[t].hi = n * (uint64_t)(t + 1) / (uint64_t)threads;
        pthread_create(&th[t], NULL, msm_worker, &jobs[t]);
    }
    jac_t acc = J_INF;
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        jadd(&acc, &acc, &jobs[t].acc);
    }
    free(jobs);
    free(th);
    fq_t x, y;
    int inf;
    to_affine(&acc, &x, &y, &inf);
    memcpy(out, x.l, 32);
    memcpy(out + 4, y.l, 32);
    return inf;
}

/* Jacobian (x, y, z Montgomery words) -> affine; returns 1 for the identity */
int g1_oracle_to_affine(uint64_t *out, const uint64_t *jac) {
    jac_t p;
    memcpy(p.x.l, jac, 32);
    memcpy(p.y.l, jac + 4, 32);
    memcpy(p.z.l, jac + 8, 32);
    fq_t x, y;
    int inf;
    to_affine(&p, &x, &y, &inf);
    memcpy(out, x.l, 32);
    memcpy(out + 4, y.l, 32);
    return inf;
}

static void gen_point(jac_t *g) {
    fq_t one = {{1, 0, 0, 0}}, two = {{2, 0, 0, 0}};
    fmul(&g->x, &one, &Q_R2); /* 1 in Montgomery form */
    fmul(&g->y, &two, &Q_R2);
    g->z = Q_ONE;
}
/* k G (k: canonical 256-bit integer, 4 words) as affine Montgomery words */
int g1_oracle_mul_gen(uint64_t *out, const uint64_t *k) {
    jac_t g, acc = J_INF;
    gen_point(&g);
    for (int bit = 255; bit >= 0; --bit) {
        jdbl(&acc, &acc);
        if ((k[bit >> 6] >> (bit & 63)) & 1) jmadd(&acc, &acc, &g.x, &g.y);
    }
    fq_t x, y;
    int inf;
    to_affine(&acc, &x, &y, &inf);
    memcpy(out, x.l, 32);
    memcpy(out + 4, y.l, 32);
    return inf;
}
/* bases[i] = (a + i d) G for i < n (a, d canonical integers below r), affine
 * Montgomery words, by repeated addition and a batched inversion */
void g1_oracle_arith_bases(uint64_t *bases, uint64_t n, const uint64_t *a, const uint64_t *d) {
    if (n == 0) return;
    uint64_t p0[8], dd[8];
    g1_oracle_mul_gen(p0, a);
    g1_oracle_mul_gen(dd, d);
    fq_t dx, dy;
    memcpy(dx.l, dd, 32);
    memcpy(dy.l, dd + 4, 32);
    jac_t *pts = (jac_t *)malloc(sizeof(jac_t) * n);
    memcpy(pts[0].x.l, p0, 32);
    memcpy(pts[0].y.l, p0 + 4, 32);
    pts[0].z = Q_ONE;
    for (uint64_t i = 1; i < n; ++i) jmadd(&pts[i], &pts[i - 1], &dx, &dy);
    /* batched inversion of the z's (none is zero for a, d chosen so that no
       a + i d vanishes mod r) */
    fq_t *pre = (fq_t *)malloc(sizeof(fq_t) * n);
    fq_t acc = Q_ONE;
    for (uint64_t i = 0; i < n; ++i) {
        pre[i] = acc;
        fmul(&acc, &acc, &pts[i].z);
    }
    fq_t inv;
    finv(&inv, &acc);
    for (uint64_t i = n; i-- > 0;) {
        fq_t zi, zi2, zi3, x, y;
        fmul(&zi, &inv, &pre[i]);
        fmul(&inv, &inv, &pts[i].z);
        fmul(&zi2, &zi, &zi);
        fmul(&zi3, &zi2, &zi);
        fmul(&x, &pts[i].x, &zi2);
        fmul(&y, &pts[i].y, &zi3);
        memcpy(bases + 8 * i, x.l, 32);
        memcpy(bases + 8 * i + 4, y.l, 32);
    }
    free(pre);
    free(pts);
}
/* Fq helpers for tests: Montgomery product, to / from Montgomery form */
void g1_oracle_fq_mul(uint64_t *o, const uint64_t *a, const uint64_t *b) {
    fq_t x, y, r;
    memcpy(x.l, a, 32);
    memcpy(y.l, b, 32);
    fmul(&r, &x, &y);
    memcpy(o, r.l, 32);
}
void g1_oracle_fq_to_mont(uint64_t *o, const uint64_t *a) {
    fq_t x, r;
    memcpy(x.l, a, 32);
    fmul(&r, &x, &Q_R2);
    memcpy(o, r.l, 32);
}
void g1_oracle_fq_from_mont(uint64_t *o, const uint64_t *a) {
    fq_t x, r, one = {{1, 0, 0, 0}};
    memcpy(x.l, a, 32);
    fmul(&r, &x, &one);
    memcpy(o, r.l, 32);
}
