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

extern "C" {
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
