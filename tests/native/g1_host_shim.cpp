// Host build of the library's Fq and G1 code (zk_gpu_amd/csrc/fq29.hpp) for
// CPU tests (tests/test_g1_host.py). Points cross the boundary as halo2
// Montgomery words: affine x, y (8 u64), Jacobian x, y, z (12 u64).
#include <string.h>

#include "../../zk_gpu_amd/csrc/fq29.hpp"

using namespace zkgpu;

static fe from64(const uint64_t* w) {
  uint32_t v[8];
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = (uint32_t)w[i];
    v[2 * i + 1] = (uint32_t)(w[i] >> 32);
  }
  return fq::q_from_words(v);
}
static void to64(uint64_t* w, const fe& a) {
  uint32_t v[8];
  fq::q_to_words(v, a);
  for (int i = 0; i < 4; ++i) w[i] = (uint64_t)v[2 * i] | ((uint64_t)v[2 * i + 1] << 32);
}
static g1j jload(const uint64_t* p) { return g1j{from64(p), from64(p + 4), from64(p + 8)}; }
static void jstore(uint64_t* o, const g1j& p) {
  to64(o, p.x);
  to64(o + 4, p.y);
  to64(o + 8, p.z);
}

// Jacobian words (X, Y, Z) -> XYZZ (X, Y, Z^2, Z^3): x = X / Z^2, y = Y / Z^3
static g1xyzz xload(const uint64_t* p) {
  const g1j j = jload(p);
  const fe zz = fq::q_mul(j.z, j.z);
  return g1xyzz{j.x, j.y, zz, fq::q_mul(zz, j.z)};
}
static void xstore(uint64_t* o, uint32_t* raw, const g1xyzz& a) {
  jstore(o, g1_xyzz_to_jacobian(a));
  if (raw) {
    memcpy(raw, a.x.l, 36);
    memcpy(raw + 9, a.y.l, 36);
    memcpy(raw + 18, a.zz.l, 36);
    memcpy(raw + 27, a.zzz.l, 36);
    raw[36] = g1_xyzz_is_identity(a) ? 1u : 0u;
  }
}

extern "C" {
// A chain of XYZZ operations on an accumulator (from the identity) and a
// saved register, with points[k] (Jacobian words) the operand of step k:
// op 0: acc = acc + p, 1: acc = p + acc, 2: acc = 2 acc (g1_xyzz_double),
// 3: acc = acc + acc, 4: saved = acc, 5: acc = acc + saved, 6: acc = saved + acc.
// out: Jacobian words of acc at the end; trace: after each step the raw
// limbs of acc (X, Y, ZZ, ZZZ) and its identity flag (37 u32).
void shim_g1_xyzz_chain(const uint64_t* points, const int* ops, int n, uint64_t* out, uint32_t* trace) {
  g1xyzz acc = g1_xyzz_identity(), saved = g1_xyzz_identity();
  for (int k = 0; k < n; ++k) {
    const g1xyzz p = xload(points + 12 * k);
    switch (ops[k]) {
      case 0: acc = g1_xyzz_add(acc, p); break;
      case 1: acc = g1_xyzz_add(p, acc); break;
      case 2: g1_xyzz_double(acc); break;
      case 3: acc = g1_xyzz_add(acc, acc); break;
      case 4: saved = acc; break;
      case 5: acc = g1_xyzz_add(acc, saved); break;
      default: acc = g1_xyzz_add(saved, acc); break;
    }
    uint64_t scratch[12];
    xstore(scratch, trace ? trace + 37 * k : nullptr, acc);
  }
  xstore(out, nullptr, acc);
}
// halo2 words (any value below 2^256) -> canonical words of 32 w mod q by
// the shift form (shift = 1) or the product form (shift = 0)
void shim_fq_from_words(const uint32_t* w, int shift, uint32_t* o) {
  fq::q_canonical_words(o, shift ? fq::q_from_words_shift(w) : fq::q_from_words(w));
}
void shim_fq_mul(const uint64_t* a, const uint64_t* b, uint64_t* o) { to64(o, fq::q_mul(from64(a), from64(b))); }
void shim_fq_add(const uint64_t* a, const uint64_t* b, uint64_t* o) { to64(o, fq::q_add(from64(a), from64(b))); }
void shim_fq_sub(const uint64_t* a, const uint64_t* b, uint64_t* o) { to64(o, fq::q_sub(from64(a), from64(b))); }
int shim_fq_is_zero(const uint64_t* a) { return fq::q_is_zero(from64(a)) ? 1 : 0; }
void shim_g1_double(const uint64_t* p, uint64_t* o) { jstore(o, g1_double(jload(p))); }
void shim_g1_add(const uint64_t* p, const uint64_t* q, uint64_t* o) { jstore(o, g1_add(jload(p), jload(q))); }
void shim_g1_add_mixed(const uint64_t* p, const uint64_t* qa, int neg, uint64_t* o) {
  g1a a;
  a.x = from64(qa);
  a.y = from64(qa + 4);
  a.inf = false;
  for (int i = 0; i < 8; ++i) a.inf = a.inf || qa[i];
  a.inf = !a.inf;
  if (neg) a = g1_neg(a);
  jstore(o, g1_add_mixed(jload(p), a));
}
// The MSM bucket sum from the identity over n affine bases (8 words each,
// x = y = 0 the identity), base i negated when neg[i]; out: Jacobian words
// (all zero for the identity). trace (optional): after each step the
// accumulator's raw limbs X, Y, ZZ, ZZZ (36 u32) and its identity flag (1 u32).
void shim_g1_bucket_sum(const uint64_t* bases, const int* neg, int n, uint64_t* out, uint32_t* trace) {
  g1xyzz acc;
  bool inf = true;
  for (int k = 0; k < n; ++k) {
    const uint64_t* qa = bases + 8 * k;
    g1a a;
    a.x = from64(qa);
    a.y = from64(qa + 4);
    bool any = false;
    for (int i = 0; i < 8; ++i) any = any || qa[i];
    a.inf = !any;
    g1_bucket_add(acc, inf, a, neg[k] != 0);
    if (trace) {
      memset(trace + 37 * k, 0, 37 * 4);
      if (!inf) {
        memcpy(trace + 37 * k, acc.x.l, 36);
        memcpy(trace + 37 * k + 9, acc.y.l, 36);
        memcpy(trace + 37 * k + 18, acc.zz.l, 36);
        memcpy(trace + 37 * k + 27, acc.zzz.l, 36);
      }
      trace[37 * k + 36] = inf ? 1u : 0u;
    }
  }
  jstore(out, inf ? g1_identity() : g1_xyzz_to_jacobian(acc));
}
}
