// BN256 base field Fq and the G1 group (y^2 = x^3 + 3) for the MSM.
//
// Field elements use the radix-2^29 form of fr29.hpp (nine limbs, 64-bit
// column accumulators, product scanning with the Montgomery reduction
// interleaved), inside the domain x * 2^261 mod q. Memory keeps halo2's
// layout: four little-endian u64 words of x * 2^256 mod q (the reference's
// BigNum in AffinePoint / ProjectivePoint, cuda/bn256.h:10-32).
//
// Every value the point formulas pass around is "small": nine normalized
// limbs (limbs 0..7 < 2^29) and value below 3q. mul keeps that (inputs below
// 3q give outputs below q (1 + 9q / 2^261) < 1.06 q); add and sub normalise and
// then reduce_value (quotient estimated from the top limb) back below 3q.
// Zero tests and stores go through the canonical form.
//
// Point formulas (Jacobian coordinates, x = X / Z^2, y = Y / Z^3, the identity
// has Z = 0) follow the same textbook formulas the reference uses
// (cuda/curve.cuh:106-297): doubling dbl-2009-l, mixed addition
// madd-2007-bl, general addition add-2007-bl, with the equal-x cases routed
// to doubling or the identity.
#pragma once
#include <stdint.h>

#include "fr29.hpp"

namespace zkgpu {
namespace fq {

// q in radix 2^29
constexpr uint32_t P29[NLIMBS] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                  0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
constexpr uint32_t NINV29 = 0x04866389u;  // -q^-1 mod 2^29
// 32-bit words of q
constexpr uint32_t P32[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                             0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
// 2^266 mod q: mul(x * 2^256, TO261) = x * 2^261
constexpr uint32_t TO261[NLIMBS] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u,
                                    0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
// 2^256 mod q: mul(x * 2^261, R256) = x * 2^256
constexpr uint32_t R256[NLIMBS] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
                                   0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
// 2^261 mod q: the one of the 2^261 domain
constexpr uint32_t ONE261[NLIMBS] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                                     0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};

ZK_HD uint32_t p29(int i) {
  return i == 0 ? P29[0] : i == 1 ? P29[1] : i == 2 ? P29[2] : i == 3 ? P29[3] : i == 4 ? P29[4]
       : i == 5 ? P29[5] : i == 6 ? P29[6] : i == 7 ? P29[7] : P29[8];
}
ZK_HD uint32_t p32(int i) {
  return i == 0 ? P32[0] : i == 1 ? P32[1] : i == 2 ? P32[2] : i == 3 ? P32[3] : i == 4 ? P32[4]
       : i == 5 ? P32[5] : i == 6 ? P32[6] : P32[7];
}

// a * b * 2^-261 mod q (not reduced); contracts as zkgpu::mont_mul.
ZK_HD fe q_mul(const fe& a, const fe& b) {
  uint32_t m[NLIMBS];
  fe o;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 2 * NLIMBS - 1; ++k) {
#pragma unroll
    for (int i = 0; i < NLIMBS; ++i) {
      const int j = k - i;
      if (j < 0 || j >= NLIMBS) continue;
      acc += (uint64_t)a.l[i] * b.l[j];
    }
#pragma unroll
    for (int i = 0; i < NLIMBS; ++i) {
      const int j = k - i;
      if (i >= k || j < 1 || j >= NLIMBS) continue;
      acc += (uint64_t)m[i] * p29(j);
    }
    if (k < NLIMBS) {
      m[k] = ((uint32_t)acc * NINV29) & LIMB_MASK;
      acc += (uint64_t)m[k] * p29(0);
    } else {
      o.l[k - NLIMBS] = (uint32_t)acc & LIMB_MASK;
    }
    acc >>= LIMB_BITS;
  }
  o.l[NLIMBS - 1] = (uint32_t)acc;
  return o;
}

// V -> V - t*q, t <= V/q estimated from the top limb; normalized in, V < 2^261
// -> normalized out, V < 3q (q's top limb equals r's, so zkgpu::QMAGIC fits).
ZK_HD void q_reduce(fe& a) {
  const uint32_t t = (uint32_t)(((uint64_t)a.l[NLIMBS - 1] * QMAGIC) >> 32);
  int64_t carry = 0;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) {
    const int64_t v = (int64_t)a.l[i] - (int64_t)t * (int64_t)p29(i) + carry;
    if (i < NLIMBS - 1) {
      a.l[i] = (uint32_t)v & LIMB_MASK;
      carry = v >> LIMB_BITS;
    } else {
      a.l[i] = (uint32_t)v;
    }
  }
}

// 4q in the borrowed limb form of zkgpu::SubConst: limbs 0..7 >= 2^29 - 1
struct Sub4 {
  static constexpr uint64_t limb_full(int i) {
    uint64_t carry = 0, v = 0;
    for (int k = 0; k <= i; ++k) {
      const uint64_t t = (uint64_t)P29[k] * 4u + carry;
      v = (k == NLIMBS - 1) ? t : (t & LIMB_MASK);
      carry = t >> LIMB_BITS;
    }
    return v;
  }
  static constexpr uint32_t k(int i) {
    return i == 0 ? (uint32_t)(limb_full(0) + (1u << LIMB_BITS))
         : i < NLIMBS - 1 ? (uint32_t)(limb_full(i) + (1u << LIMB_BITS) - 1u)
         : (uint32_t)(limb_full(i) - 1u);
  }
};

ZK_HD fe q_add(const fe& a, const fe& b) {
  fe o = zkgpu::add(a, b);
  normalize(o);
  q_reduce(o);
  return o;
}
// a - b (b normalized, V_b < 3q)
ZK_HD fe q_sub(const fe& a, const fe& b) {
  fe o;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) o.l[i] = a.l[i] + (Sub4::k(i) - b.l[i]);
  normalize(o);
  q_reduce(o);
  return o;
}
ZK_HD fe q_dbl(const fe& a) { return q_add(a, a); }
ZK_HD fe q_sqr(const fe& a) { return q_mul(a, a); }

// words (value < 2^256): subtract q once when the value is >= q
ZK_HD void cond_sub_q_words(uint32_t w[8]) {
  uint32_t t[8];
  uint32_t bw = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t d = (uint64_t)w[i] - p32(i) - bw;
    t[i] = (uint32_t)d;
    bw = (uint32_t)(d >> 63);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = bw ? w[i] : t[i];
}
// small a -> canonical words of a mod q
ZK_HD void q_canonical_words(uint32_t w[8], fe a) {
  q_reduce(a);
  pack_words(w, a);  // V < 3q < 2^256
  cond_sub_q_words(w);
  cond_sub_q_words(w);
}
ZK_HD bool q_is_zero(const fe& a) {
  uint32_t w[8];
  q_canonical_words(w, a);
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) o |= w[i];
  return o == 0;
}
ZK_HD bool q_equal(const fe& a, const fe& b) { return q_is_zero(q_sub(a, b)); }

ZK_HD fe q_from_const(const uint32_t c[NLIMBS]) {
  fe a;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) a.l[i] = c[i];
  return a;
}
ZK_HD fe q_zero() {
  fe a;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) a.l[i] = 0;
  return a;
}
// halo2 words (x * 2^256 mod q, < 2^256) -> small x * 2^261
ZK_HD fe q_from_words(const uint32_t w[8]) { return q_mul(unpack_words(w), q_from_const(TO261)); }
// small x * 2^261 -> canonical halo2 words x * 2^256 mod q
ZK_HD void q_to_words(uint32_t w[8], const fe& a) { q_canonical_words(w, q_mul(a, q_from_const(R256))); }

}  // namespace fq

// G1 point in Jacobian coordinates over Fq (2^261 domain); Z = 0 is the identity.
struct g1j {
  fe x, y, z;
};
// affine point; inf = the identity (halo2 stores it as x = y = 0)
struct g1a {
  fe x, y;
  bool inf;
};

ZK_HD g1j g1_identity() { return g1j{fq::q_zero(), fq::q_zero(), fq::q_zero()}; }
ZK_HD bool g1_is_identity(const g1j& p) { return fq::q_is_zero(p.z); }
ZK_HD g1j g1_from_affine(const g1a& a) {
  if (a.inf) return g1_identity();
  return g1j{a.x, a.y, fq::q_from_const(fq::ONE261)};
}
ZK_HD g1a g1_neg(const g1a& a) {
  g1a o = a;
  o.y = fq::q_sub(fq::q_zero(), a.y);
  return o;
}

// 2P, dbl-2009-l (a = 0): 2M + 5S
ZK_HD g1j g1_double(const g1j& p) {
  using namespace fq;
  if (g1_is_identity(p)) return p;
  const fe A = q_sqr(p.x), B = q_sqr(p.y), C = q_sqr(B);
  const fe t = q_sqr(q_add(p.x, B));
  const fe D = q_dbl(q_sub(q_sub(t, A), C));
  const fe E = q_add(q_dbl(A), A);
  const fe F = q_sqr(E);
  g1j o;
  o.x = q_sub(F, q_dbl(D));
  const fe C8 = q_dbl(q_dbl(q_dbl(C)));
  o.y = q_sub(q_mul(E, q_sub(D, o.x)), C8);
  o.z = q_dbl(q_mul(p.y, p.z));
  return o;
}

// P + Q with Q affine, madd-2007-bl: 7M + 4S
ZK_HD g1j g1_add_mixed(const g1j& p, const g1a& q) {
  using namespace fq;
  if (q.inf) return p;
  if (g1_is_identity(p)) return g1_from_affine(q);
  const fe Z1Z1 = q_sqr(p.z);
  const fe U2 = q_mul(q.x, Z1Z1);
  const fe S2 = q_mul(q.y, q_mul(p.z, Z1Z1));
  const fe H = q_sub(U2, p.x);
  const fe R = q_dbl(q_sub(S2, p.y));
  if (q_is_zero(H)) {
    if (q_is_zero(R)) return g1_double(p);
    return g1_identity();
  }
  const fe HH = q_sqr(H);
  const fe I = q_dbl(q_dbl(HH));
  const fe J = q_mul(H, I);
  const fe V = q_mul(p.x, I);
  g1j o;
  o.x = q_sub(q_sub(q_sqr(R), J), q_dbl(V));
  o.y = q_sub(q_mul(R, q_sub(V, o.x)), q_dbl(q_mul(p.y, J)));
  o.z = q_sub(q_sub(q_sqr(q_add(p.z, H)), Z1Z1), HH);
  return o;
}

// P + Q, add-2007-bl: 11M + 5S
ZK_HD g1j g1_add(const g1j& p, const g1j& q) {
  using namespace fq;
  if (g1_is_identity(p)) return q;
  if (g1_is_identity(q)) return p;
  const fe Z1Z1 = q_sqr(p.z), Z2Z2 = q_sqr(q.z);
  const fe U1 = q_mul(p.x, Z2Z2), U2 = q_mul(q.x, Z1Z1);
  const fe S1 = q_mul(p.y, q_mul(q.z, Z2Z2));
  const fe S2 = q_mul(q.y, q_mul(p.z, Z1Z1));
  const fe H = q_sub(U2, U1);
  const fe R = q_dbl(q_sub(S2, S1));
  if (q_is_zero(H)) {
    if (q_is_zero(R)) return g1_double(p);
    return g1_identity();
  }
  const fe I = q_sqr(q_dbl(H));
  const fe J = q_mul(H, I);
  const fe V = q_mul(U1, I);
  g1j o;
  o.x = q_sub(q_sub(q_sqr(R), J), q_dbl(V));
  o.y = q_sub(q_mul(R, q_sub(V, o.x)), q_dbl(q_mul(S1, J)));
  o.z = q_mul(q_sub(q_sub(q_sqr(q_add(p.z, q.z)), Z1Z1), Z2Z2), H);
  return o;
}

}  // namespace zkgpu
