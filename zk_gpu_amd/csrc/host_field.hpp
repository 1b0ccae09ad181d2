// Host-side BN256 Fr helpers used to derive per-call constants (roots of
// unity, inverses, n^-1, the 2^261-domain form of scales). Four 64-bit words,
// Montgomery form with R = 2^256 (the FFI layout, reference
// cuda/bn256.h:10-13); CIOS multiply over unsigned __int128.
#pragma once
#include <stdint.h>

namespace zkgpu {
namespace host {

struct Fr {
  uint64_t w[4];
};

constexpr Fr MOD = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
constexpr Fr ONE = {{0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL}};
constexpr Fr R2 = {{0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}};
constexpr uint64_t NINV = 0xc2e1f593efffffffULL;
// 7^((r-1)/2^28), the 2^28-th root of unity derived from the generator 7
constexpr Fr ROOT28 = {{0x9632c7c5b639feb8ULL, 0x985ce3400d0ff299ULL, 0xb2dd880001b0ecd8ULL, 0x1d69070d6d98ce29ULL}};
constexpr uint32_t TWO_ADICITY = 28;

inline bool geq(const Fr& a, const Fr& b) {
  for (int i = 3; i >= 0; --i) {
    if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
  }
  return true;
}
inline bool is_canonical(const Fr& a) { return !geq(a, MOD); }
inline void sub_mod_raw(Fr& a, const Fr& b) {
  unsigned __int128 bw = 0;
  for (int i = 0; i < 4; ++i) {
    unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - bw;
    a.w[i] = (uint64_t)d;
    bw = (d >> 64) & 1;
  }
}
inline Fr mul(const Fr& a, const Fr& b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < 4; ++j) {
      unsigned __int128 p = (unsigned __int128)a.w[i] * b.w[j] + t[j] + c;
      t[j] = (uint64_t)p;
      c = (uint64_t)(p >> 64);
    }
    unsigned __int128 s = (unsigned __int128)t[4] + c;
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    const uint64_t m = t[0] * NINV;
    unsigned __int128 p = (unsigned __int128)m * MOD.w[0] + t[0];
    c = (uint64_t)(p >> 64);
    for (int j = 1; j < 4; ++j) {
      p = (unsigned __int128)m * MOD.w[j] + t[j] + c;
      t[j - 1] = (uint64_t)p;
      c = (uint64_t)(p >> 64);
    }
    s = (unsigned __int128)t[4] + c;
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  Fr o = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || geq(o, MOD)) sub_mod_raw(o, MOD);
  return o;
}
inline Fr pow(const Fr& b, const Fr& e_plain) {
  Fr acc = ONE;
  for (int i = 3; i >= 0; --i)
    for (int k = 63; k >= 0; --k) {
      acc = mul(acc, acc);
      if ((e_plain.w[i] >> k) & 1) acc = mul(acc, b);
    }
  return acc;
}
inline Fr pow_u64(const Fr& b, uint64_t e) {
  Fr ee = {{e, 0, 0, 0}};
  return pow(b, ee);
}
inline Fr inv(const Fr& a) {
  Fr e = MOD;
  e.w[0] -= 2;
  return pow(a, e);
}
// canonical integer x (< 2^64) -> Montgomery form
inline Fr from_u64(uint64_t x) {
  Fr v = {{x, 0, 0, 0}};
  return mul(v, R2);
}
inline Fr root_of_unity(uint32_t log_n) {
  Fr w = ROOT28;
  for (uint32_t i = log_n; i < TWO_ADICITY; ++i) w = mul(w, w);
  return w;
}
inline bool equal(const Fr& a, const Fr& b) {
  return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.w[2] == b.w[2] && a.w[3] == b.w[3];
}

// x * 2^256 (Montgomery words) -> x * 2^261 mod r as nine normalized 29-bit limbs.
inline void to_limbs261(uint32_t out[9], const Fr& a) {
  // value * 32 mod r: shift into five words, then subtract 16r, 8r, ..., r when possible
  uint64_t v[5] = {a.w[0] << 5, (a.w[1] << 5) | (a.w[0] >> 59), (a.w[2] << 5) | (a.w[1] >> 59),
                   (a.w[3] << 5) | (a.w[2] >> 59), a.w[3] >> 59};
  for (int s = 4; s >= 0; --s) {
    uint64_t m[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      m[i] |= MOD.w[i] << s;
      if (s && i + 1 < 5) m[i + 1] |= MOD.w[i] >> (64 - s);
    }
    // v >= m ?
    bool ge = true;
    for (int i = 4; i >= 0; --i) {
      if (v[i] != m[i]) { ge = v[i] > m[i]; break; }
    }
    if (ge) {
      unsigned __int128 bw = 0;
      for (int i = 0; i < 5; ++i) {
        unsigned __int128 d = (unsigned __int128)v[i] - m[i] - bw;
        v[i] = (uint64_t)d;
        bw = (d >> 64) & 1;
      }
    }
  }
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, q = bit / 64, s = bit % 64;
    uint64_t x = v[q] >> s;
    if (s > 35 && q + 1 < 5) x |= v[q + 1] << (64 - s);
    out[i] = (uint32_t)(x & (i < 8 ? 0x1fffffffULL : 0xffffffffULL));
  }
}

inline void to_words32(uint32_t out[8], const Fr& a) {
  for (int i = 0; i < 4; ++i) {
    out[2 * i] = (uint32_t)a.w[i];
    out[2 * i + 1] = (uint32_t)(a.w[i] >> 32);
  }
}
inline Fr from_ptr(const uint64_t* p) { return Fr{{p[0], p[1], p[2], p[3]}}; }

}  // namespace host
}  // namespace zkgpu
