// BN256 scalar field Fr in a redundant radix-2^29 representation for gfx950.
//
// Why 29-bit limbs: every product of two limbs is < 2^58, so a column of up
// to 18 partial products plus the carry fits a 64-bit accumulator without
// overflow. The Montgomery multiply is then a plain chain of
// v_mad_u64_u32 (full rate on gfx950) with no carry-flag traffic: 162
// multiply-adds and about 60 shift/mask/multiply-low operations, no VCC
// hazards, and ordinary C that the compiler schedules freely.
//
// Element: value = sum_i l[i] * 2^(29 i), i < 9 (261 bits of capacity).
// "Normalized" means l[0..7] < 2^29 (l[8] holds the rest). Values are kept
// congruent mod r but not reduced: bounds are tracked by the callers.
//
// Data in memory keeps halo2's layout: four little-endian u64 words of
// x * 2^256 mod r (reference cuda/bn256.h:10-13). Constants (twiddles,
// scales) are kept as c * 2^261 mod r, so mont_mul(x * 2^256, c * 2^261)
// = x * c * 2^256: multiplying by a constant never leaves the data domain.
//
// Contracts (V = value, units of r):
//   mont_mul(a, b): a limbs <= MUL_A_LIMB_MAX, b normalized with V < r.
//                   out normalized, V_out < r * (1 + V_a / 2^261) < r * (1 + V_a / 169).
//   add(a, b):      limb-wise, no carries.
//   sub<C>(a, b):   a + C*r - b; b normalized with V_b < (C-1) r.
//                   out limbs l[i] <= a[i] + 2^30 (i < 8).
//   normalize(a):   carry propagation; value unchanged.
//   reduce_value(a): a normalized, V < 2^261 -> normalized, V < 3r.
//   to_canonical_words(a): limbs 0..7 <= MUL_A_LIMB_MAX, V < 2^261 -> canonical u32[8] words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZK_HD __host__ __device__ __forceinline__

namespace zkgpu {

constexpr uint32_t LIMB_BITS = 29;
constexpr uint32_t LIMB_MASK = (1u << LIMB_BITS) - 1u;
constexpr int NLIMBS = 9;
// largest limb of the multiplier's first operand for which no column can
// overflow 64 bits (9 * 0xC0000000 * 2^29 + 9 * 2^29 * 2^29 + carry < 2^64)
constexpr uint32_t MUL_A_LIMB_MAX = 0xC0000000u;

struct fe {
  uint32_t l[NLIMBS];
};

// r in radix 2^29
constexpr uint32_t P29[NLIMBS] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                                  0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
// -r^-1 mod 2^29
constexpr uint32_t NINV29 = 0x0fffffffu;
// 32-bit words of r (for canonicalisation on packed words)
constexpr uint32_t P32[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                             0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};

ZK_HD uint32_t p29(int i) {
  return i == 0 ? P29[0] : i == 1 ? P29[1] : i == 2 ? P29[2] : i == 3 ? P29[3] : i == 4 ? P29[4]
       : i == 5 ? P29[5] : i == 6 ? P29[6] : i == 7 ? P29[7] : P29[8];
}
ZK_HD uint32_t p32(int i) {
  return i == 0 ? P32[0] : i == 1 ? P32[1] : i == 2 ? P32[2] : i == 3 ? P32[3] : i == 4 ? P32[4]
       : i == 5 ? P32[5] : i == 6 ? P32[6] : P32[7];
}

// r for the column-asm multiplier
struct FrMod {
  static constexpr uint32_t ninv = NINV29;
  static constexpr uint32_t p(int i) { return P29[i]; }
};

}  // namespace zkgpu

#ifdef __HIP_DEVICE_COMPILE__
#include "mont_cols.hpp"
#endif

namespace zkgpu {

// Montgomery product a * b * 2^-261 mod r (not reduced), product scanning
// with the reduction interleaved per column (FIPS, Koc-Acar-Kaliski 1996).
// CHAIN = false: the compiler splits each column into chains that start from
// zero and adds the carry last (more independent work per wave, one 64-bit
// add more per column). CHAIN = true (device): each column is one chain of
// multiply-adds from the carry, written as one asm statement per column
// (mont_cols.hpp): fewer instructions and registers, longer dependences.
// The host build computes the same product in plain C either way.
template <bool CHAIN = false>
ZK_HD fe mont_mul(const fe& a, const fe& b) {
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (CHAIN) return mont_mul_cols<FrMod>(a, b);
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

// mont_mul with b the same in every lane of the wave (a root of a register
// DFT, a uniform scale): the single-chain form takes b from SGPRs.
template <bool CHAIN = false>
ZK_HD fe mont_mul_u(const fe& a, const fe& b) {
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (CHAIN) return mont_mul_cols<FrMod, true>(a, b);
#endif
  return mont_mul<false>(a, b);
}

ZK_HD fe add(const fe& a, const fe& b) {
  fe o;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) o.l[i] = a.l[i] + b.l[i];
  return o;
}

// Limbs of K = C*r written so that K[i] >= 2^29 - 1 for i < 8 (borrowing one
// unit from the limb above): K[0] = s0 + 2^29, K[i] = s_i + 2^29 - 1,
// K[8] = s8 - 1, where s are the standard radix-2^29 limbs of C*r.
template <uint32_t C>
struct SubConst {
  static constexpr uint64_t limb_full(int i) {
    // standard limb i of C*r, with the top limb unmasked
    uint64_t carry = 0, v = 0;
    for (int q = 0; q <= i; ++q) {
      const uint64_t t = (uint64_t)P29[q] * C + carry;
      v = (q == NLIMBS - 1) ? t : (t & LIMB_MASK);
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

template <uint32_t C>
ZK_HD fe sub(const fe& a, const fe& b) {
  fe o;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) o.l[i] = a.l[i] + (SubConst<C>::k(i) - b.l[i]);
  return o;
}

ZK_HD void normalize(fe& a) {
#pragma unroll
  for (int i = 0; i < NLIMBS - 1; ++i) {
    a.l[i + 1] += a.l[i] >> LIMB_BITS;
    a.l[i] &= LIMB_MASK;
  }
}

// floor(2^32 / (r8 + 1)) where r8 = P29[8]
constexpr uint32_t QMAGIC = (uint32_t)((1ull << 32) / (0x0030644eull + 1ull));

// V -> V - q*r with q <= V/r estimated from the top limb; result V < 3r.
ZK_HD void reduce_value(fe& a) {
  const uint32_t q = (uint32_t)(((uint64_t)a.l[NLIMBS - 1] * QMAGIC) >> 32);
  int64_t carry = 0;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) {
    const int64_t t = (int64_t)a.l[i] - (int64_t)q * (int64_t)p29(i) + carry;
    if (i < NLIMBS - 1) {
      a.l[i] = (uint32_t)t & LIMB_MASK;
      carry = t >> LIMB_BITS;  // arithmetic shift
    } else {
      a.l[i] = (uint32_t)t;
    }
  }
}

// bit offset helpers for packing 9 x 29 <-> 8 x 32
ZK_HD void pack_words(uint32_t w[8], const fe& a) {
  // normalized a: the words of V mod 2^256
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = (32 * q) / 29, s = 32 * q - 29 * i;
    uint32_t v = a.l[i] >> s;
    if (29 - s < 32 && i + 1 < NLIMBS) v |= a.l[i + 1] << (29 - s);
    if (2 * 29 - s < 32 && i + 2 < NLIMBS) v |= a.l[i + 2] << (2 * 29 - s);
    w[q] = v;
  }
}
ZK_HD fe unpack_words(const uint32_t w[8]) {
  fe a;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) {
    const int q = (29 * i) / 32, s = 29 * i - 32 * q;
    uint32_t v = w[q] >> s;
    if (s > 3 && q + 1 < 8) v |= w[q + 1] << (32 - s);
    a.l[i] = (i < NLIMBS - 1) ? (v & LIMB_MASK) : v;
  }
  return a;
}

// a - b - borrow_in on 32-bit words, borrow out in {0, 1}. On the device a
// chain of these is one v_sub_co_u32 / v_subb_co_u32 sequence through VCC.
ZK_HD uint32_t sub_borrow(uint32_t a, uint32_t b, uint32_t bin, uint32_t* bout) {
#if defined(__clang__)
  unsigned int o;
  const uint32_t d = __builtin_subc(a, b, bin, &o);
  *bout = o;
  return d;
#else
  const uint64_t t = (uint64_t)a - b - bin;
  *bout = (uint32_t)(t >> 63);
  return (uint32_t)t;
#endif
}

// words (value < 2^256): subtract r once if the value is >= r
ZK_HD void cond_sub_r_words(uint32_t w[8]) {
  uint32_t t[8];
  uint32_t bw = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) t[i] = sub_borrow(w[i], p32(i), bw, &bw);
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = bw ? w[i] : t[i];
}

// a with limbs 0..7 <= MUL_A_LIMB_MAX (the multiplier's operand bound, so
// normalisation cannot overflow a limb) and V < 2^261 -> canonical words of
// V mod r.
// After normalisation the top limb is L = floor(V / 2^232) < 2^29, and
// q = floor(L * QMAGIC / 2^32) satisfies q <= L / (r8 + 1) <= V / r and
// L - q r8 < L / (r8 + 1) + L r8 / 2^32 + r8 < 1.13 r8 (r8 = r's top limb),
// so 0 <= V - q r < 1.13 r + 2^232 < 2r. As 2r < 2^256 the subtraction is
// exact on the low 256 bits: pack them, subtract q r mod 2^256 (one
// multiply-add per word and a borrow chain), then subtract r once if needed.
ZK_HD void to_canonical_words(uint32_t w[8], fe a) {
  normalize(a);
  const uint32_t q = (uint32_t)(((uint64_t)a.l[NLIMBS - 1] * QMAGIC) >> 32);
  pack_words(w, a);  // V mod 2^256
  uint32_t c = 0, bw = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t t = (uint64_t)q * p32(i) + c;
    c = (uint32_t)(t >> 32);
    w[i] = sub_borrow(w[i], (uint32_t)t, bw, &bw);
  }
  cond_sub_r_words(w);
}

// 2^266 mod r in radix 2^29: mont_mul(x * 2^256, TO261) = x * 2^261
constexpr uint32_t TO261[NLIMBS] = {0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau,
                                    0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
// 2^517 mod r: mont_mul(x, ENCODE256) = x * 2^256 (canonical integer -> halo2 Montgomery form)
constexpr uint32_t ENCODE256[NLIMBS] = {0x142db4dfu, 0x19d6990eu, 0x1472f48cu, 0x06dbe7e3u, 0x0b84d579u,
                                        0x10f9faf7u, 0x121f4380u, 0x17a112deu, 0x001275c7u};
// 2^261 mod r (the constant 1 in the 2^261 domain)
constexpr uint32_t ONE261[NLIMBS] = {0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu,
                                     0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};

ZK_HD fe fe_from_const(const uint32_t c[NLIMBS]) {
  fe a;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) a.l[i] = c[i];
  return a;
}

}  // namespace zkgpu
