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
// to doubling or the identity. The MSM works in XYZZ coordinates instead
// (x = X / ZZ, y = Y / ZZZ): mixed additions into the bucket sums
// (madd-2008-s, g1_bucket_add), full additions and doublings of bucket,
// segment, run and window sums (add-2008-s, dbl-2008-s-1); the Jacobian
// formulas serve the synthetic bases and the host-side tests.
#pragma once
#include <stdint.h>

#include "fr29.hpp"

namespace zkgpu {
namespace fq {

// q in radix 2^29
constexpr uint32_t P29[NLIMBS] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                  0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
constexpr uint32_t NINV29 = 0x04866389u;  // -q^-1 mod 2^29
// q for the column-asm multiplier (mont_cols.hpp, device code)
struct FqMod {
  static constexpr uint32_t ninv = NINV29;
  static constexpr uint32_t p(int i) { return P29[i]; }
};
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

// a * b * 2^-261 mod q (not reduced); contracts as zkgpu::mont_mul. On the
// device each column is one chain of multiply-adds from the carry
// (mont_mul_cols); the host computes the same product in plain C.
ZK_HD fe q_mul(const fe& a, const fe& b) {
#ifdef __HIP_DEVICE_COMPILE__
  return mont_mul_cols<FqMod>(a, b);
#endif
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
  for (int i = 0; i < 8; ++i) t[i] = sub_borrow(w[i], p32(i), bw, &bw);
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
// The same residue, x * 2^261 = 32 (x * 2^256) mod q, by a 5-bit shift of
// the limbs (value below 2^261) and one value reduction instead of a
// product: normalized out, value below 3q.
ZK_HD fe q_from_words_shift(const uint32_t w[8]) {
  fe a = unpack_words(w);  // limbs 0..7 < 2^29, top limb < 2^24
  a.l[NLIMBS - 1] = (a.l[NLIMBS - 1] << 5) | (a.l[NLIMBS - 2] >> (LIMB_BITS - 5));
#pragma unroll
  for (int i = NLIMBS - 2; i >= 1; --i) a.l[i] = ((a.l[i] << 5) | (a.l[i - 1] >> (LIMB_BITS - 5))) & LIMB_MASK;
  a.l[0] = (a.l[0] << 5) & LIMB_MASK;
  q_reduce(a);
  return a;
}
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

namespace fq {
// C q in the borrowed limb form of zkgpu::SubConst (limbs 0..7 >= 2^29 - 1)
template <uint32_t C>
struct SubQ {
  static constexpr uint64_t limb_full(int i) {
    uint64_t carry = 0, v = 0;
    for (int k = 0; k <= i; ++k) {
      const uint64_t t = (uint64_t)P29[k] * C + carry;
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
// a + C q - b, no carries: b normalized with V_b < (C - 1) q; limbs 0..7 of
// the result <= a[i] + 2^30, value V_a + C - V_b (units of q).
template <uint32_t C>
ZK_HD fe q_sub_lazy(const fe& a, const fe& b) {
  fe o;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) o.l[i] = a.l[i] + (SubQ<C>::k(i) - b.l[i]);
  return o;
}
ZK_HD fe q_norm(fe a) {
  normalize(a);
  return a;
}
// a normalized with V < 3q: a == 0 mod q, i.e. a is 0, q or 2q limb for limb
ZK_HD bool q_is_zero3(const fe& a) {
  uint32_t z0 = 0, z1 = 0, z2 = 0;
  uint64_t c2 = 0;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) {
    const uint64_t t = 2ull * P29[i] + c2;  // limbs of 2q, carried at compile time
    const uint32_t q2 = (i < NLIMBS - 1) ? (uint32_t)(t & LIMB_MASK) : (uint32_t)t;
    c2 = t >> LIMB_BITS;
    z0 |= a.l[i];
    z1 |= a.l[i] ^ P29[i];
    z2 |= a.l[i] ^ q2;
  }
  return z0 == 0 || z1 == 0 || z2 == 0;
}
}  // namespace fq

// The MSM's bucket accumulator in XYZZ coordinates: x = X / ZZ,
// y = Y / ZZZ with ZZ^3 = ZZZ^2 (ZZ = Z^2, ZZZ = Z^3 of a Jacobian Z). A
// mixed addition costs 8M + 2S there (madd-2008-s), against 7M + 4S in
// Jacobian coordinates.
struct g1xyzz {
  fe x, y, zz, zzz;
};

// Jacobian form of a (not the identity): (X ZZ^2, Y ZZZ^2, ZZZ), since
// ZZZ^2 = ZZ^3 gives X ZZ^2 / ZZZ^2 = X / ZZ and Y ZZZ^2 / ZZZ^3 = Y / ZZZ.
ZK_HD g1j g1_xyzz_to_jacobian(const g1xyzz& a) {
  using namespace fq;
  return g1j{q_mul(a.x, q_mul(a.zz, a.zz)), q_mul(a.y, q_mul(a.zzz, a.zzz)), a.zzz};
}

// 2A in XYZZ coordinates, dbl-2008-s-1 (a = 0): 6M + 3S. Bounds in units of
// q as for g1_bucket_add: U = 2 Y1 < 6, V = U^2 < 1.22, W = U V < 1.05,
// S = X1 V < 1.03, M = 3 X1^2 < 3.2, M^2 < 1.07, X3 = M^2 + 6q - 2S < 7.07,
// S + 9q - X3 < 10.03, M (S - X3) < 1.19, W Y1 < 1.02, Y3 < 4.19; X3 and Y3
// value-reduced below 3q. BN256 G1 has no point of order two, so Y1 != 0.
ZK_HD void g1_xyzz_double(g1xyzz& a) {
  using namespace fq;
  const fe U = q_norm(zkgpu::add(a.y, a.y));
  const fe V = q_mul(U, U);
  const fe W = q_mul(U, V);
  const fe S = q_mul(a.x, V);
  const fe XX = q_mul(a.x, a.x);
  const fe M = q_norm(zkgpu::add(zkgpu::add(XX, XX), XX));
  fe X3 = q_norm(q_sub_lazy<3>(q_sub_lazy<3>(q_mul(M, M), S), S));
  fe Y3 = q_norm(q_sub_lazy<3>(q_mul(q_sub_lazy<9>(S, X3), M), q_mul(W, a.y)));
  a.zz = q_mul(V, a.zz);
  a.zzz = q_mul(W, a.zzz);
  q_reduce(X3);
  q_reduce(Y3);
  a.x = X3;
  a.y = Y3;
}

// The MSM's bucket sum: acc += (neg ? -q : q), with the identity carried in
// *inf (acc is meaningful only when !*inf). Invariant between calls: X, Y
// normalized with value < 3q; ZZ, ZZZ normalized products (value < 1.07q).
// madd-2008-s with lazy field steps; bounds in units of q, from mul's
// V_out < 1 + V_a V_b q / 2^261 with q / 2^261 < 1/169 (operand a: limbs
// <= MUL_A_LIMB_MAX; b: normalized):
//   y2 (q.y or 3q - q.y, limbs < 2^30)  < 3
//   U2 = x2 ZZ1 < 1.01, S2 = y2 ZZZ1 < 1.02
//   P = U2 + 4q - X1 (normalized) < 5.01;  R = S2 + 4q - Y1 (normalized) < 5.02
//   PP < 1.15, PPP = P PP < 1.04, Q = X1 PP < 1.03, R^2 < 1.15
//   X3 = R^2 + 3q - PPP + 3q - Q + 3q - Q (normalized) < 10.15
//   Q + 12q - X3 (limbs < 2^29 + 2^30) < 13.03, R (Q - X3) < 1.39, Y1 PPP < 1.02
//   Y3 = R (Q - X3) + 3q - Y1 PPP < 4.39;  ZZ3 = ZZ1 PP, ZZZ3 = ZZZ1 PPP < 1.01
// X3 and Y3 are normalized and value-reduced (below 3q). ZZ1 != 0, so ZZ3
// == 0 exactly when P == 0: then the base equals +-acc, and the sum is 2 acc
// (R == 0) or the identity.
ZK_HD void g1_bucket_add(g1xyzz& acc, bool& inf, const g1a& q, bool neg) {
  using namespace fq;
  if (q.inf) return;
  const fe yn = q_sub_lazy<3>(q_zero(), q.y);
  fe y2;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) y2.l[i] = neg ? yn.l[i] : q.y.l[i];
  if (inf) {
    acc.x = q.x;
    acc.y = q_norm(y2);
    acc.zz = q_from_const(fq::ONE261);
    acc.zzz = acc.zz;
    inf = false;
    return;
  }
  const fe U2 = q_mul(q.x, acc.zz);
  const fe S2 = q_mul(y2, acc.zzz);
  const fe P = q_norm(q_sub_lazy<4>(U2, acc.x));
  const fe R = q_norm(q_sub_lazy<4>(S2, acc.y));
  const fe PP = q_mul(P, P);
  const fe PPP = q_mul(P, PP);
  const fe ZZ3 = q_mul(acc.zz, PP);
  if (q_is_zero3(ZZ3)) {
    if (q_is_zero(R)) g1_xyzz_double(acc);
    else inf = true;
    return;
  }
  const fe Q = q_mul(acc.x, PP);
  fe X3 = q_norm(q_sub_lazy<3>(q_sub_lazy<3>(q_sub_lazy<3>(q_mul(R, R), PPP), Q), Q));
  fe Y3 = q_norm(q_sub_lazy<3>(q_mul(q_sub_lazy<12>(Q, X3), R), q_mul(acc.y, PPP)));
  acc.zzz = q_mul(acc.zzz, PPP);
  acc.zz = ZZ3;
  q_reduce(X3);
  q_reduce(Y3);
  acc.x = X3;
  acc.y = Y3;
}

ZK_HD g1xyzz g1_xyzz_identity() { return g1xyzz{fq::q_zero(), fq::q_zero(), fq::q_zero(), fq::q_zero()}; }
// the identity has ZZ = 0 (and ZZZ = 0); every other point has ZZ != 0 mod q
// (ZZ is always a normalized product or constant below 3q)
ZK_HD bool g1_xyzz_is_identity(const g1xyzz& a) { return fq::q_is_zero3(a.zz); }

// P1 + P2 in XYZZ coordinates, add-2008-s: 12M + 2S, with lazy field steps.
// Inputs and outputs: X, Y normalized below 3q; ZZ, ZZZ normalized products
// below 1.1q. Bounds in units of q, as for g1_bucket_add:
//   U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1 < 1.02
//   P = U2 + 3q - U1 (normalized) < 4.02;  R = S2 + 3q - S1 (normalized) < 4.02
//   PP < 1.1, PPP = P PP < 1.03, Q = U1 PP < 1.01, R^2 < 1.1, ZZ1 ZZ2 < 1.01
//   X3 = R^2 + 3q - PPP + 3q - Q + 3q - Q (normalized) < 10.1
//   Q + 12q - X3 < 13.01, R (Q - X3) < 1.31, S1 PPP < 1.01, Y3 < 4.31
//   ZZ3 = ZZ1 ZZ2 PP, ZZZ3 = ZZZ1 ZZZ2 PPP < 1.01
// X3 and Y3 are value-reduced below 3q. ZZ1 ZZ2 != 0, so ZZ3 == 0 exactly
// when P == 0, i.e. x1 == x2: then the sum is 2 P1 (R == 0) or the identity.
ZK_HD g1xyzz g1_xyzz_add(const g1xyzz& p, const g1xyzz& q) {
  using namespace fq;
  if (g1_xyzz_is_identity(p)) return q;
  if (g1_xyzz_is_identity(q)) return p;
  const fe U1 = q_mul(p.x, q.zz), U2 = q_mul(q.x, p.zz);
  const fe S1 = q_mul(p.y, q.zzz), S2 = q_mul(q.y, p.zzz);
  const fe P = q_norm(q_sub_lazy<3>(U2, U1)), R = q_norm(q_sub_lazy<3>(S2, S1));
  const fe PP = q_mul(P, P);
  const fe ZZ3 = q_mul(q_mul(p.zz, q.zz), PP);
  if (q_is_zero3(ZZ3)) {
    if (!q_is_zero(R)) return g1_xyzz_identity();
    g1xyzz d = p;
    g1_xyzz_double(d);
    return d;
  }
  const fe PPP = q_mul(P, PP), Q = q_mul(U1, PP);
  fe X3 = q_norm(q_sub_lazy<3>(q_sub_lazy<3>(q_sub_lazy<3>(q_mul(R, R), PPP), Q), Q));
  fe Y3 = q_norm(q_sub_lazy<3>(q_mul(q_sub_lazy<12>(Q, X3), R), q_mul(S1, PPP)));
  g1xyzz o;
  o.zz = ZZ3;
  o.zzz = q_mul(q_mul(p.zzz, q.zzz), PPP);
  q_reduce(X3);
  q_reduce(Y3);
  o.x = X3;
  o.y = Y3;
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
