// Radix-R decimation-in-frequency networks in registers (R = 2, 4, 8) with
// the compile-time normalisation schedule that keeps every operation inside
// the limb and value contracts of fr29.hpp. Host and device: the pass kernels
// run them, and a host build of the same code is driven at worst-case bounds
// by the CPU tests.
#pragma once
#include <stdint.h>

#include <type_traits>

#include "fr29.hpp"

namespace zkgpu {

// unpacked table entry: 9 limbs in 12 u32 (48 bytes)
ZK_HD fe load_limbs48(const void* base, uint64_t idx) {
  const uint4* p = (const uint4*)base + 3 * idx;
  const uint4 a = p[0], b = p[1], c = p[2];
  fe r;
  r.l[0] = a.x; r.l[1] = a.y; r.l[2] = a.z; r.l[3] = a.w;
  r.l[4] = b.x; r.l[5] = b.y; r.l[6] = b.z; r.l[7] = b.w;
  r.l[8] = c.x;
  return r;
}

template <int R>
struct Log2 { static constexpr int value = R <= 1 ? 0 : 1 + Log2<R / 2>::value; };

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, N). Keeps
// register-array indices constant without relying on the unroller.
template <int N, int B = 0, class F>
ZK_HD void static_for(F&& f) {
  if constexpr (B < N) {
    f(std::integral_constant<int, B>{});
    static_for<N, B + 1>(f);
  }
}

__host__ __device__ constexpr int bitrev_c(int v, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) { r = (r << 1) | (v & 1); v >>= 1; }
  return r;
}

// ---------------------------------------------------------------- DFT

// Bound state of one register: lo = max of limbs 0..7, top = max of limb 8,
// v = value bound in units of r (value < v * r), norm = limbs 0..7 < 2^29.
struct Bnd {
  uint64_t lo, top;
  uint32_t v;
  bool norm;
};
constexpr uint64_t LIMB_MAX32 = 0xFFFFFFFFull;
constexpr uint64_t MUL_IN_MAX = MUL_A_LIMB_MAX;   // multiplier operand limb bound
constexpr uint32_t MUL_IN_VMAX = 160;             // keeps multiplier outputs below 2r
constexpr uint64_t TOP_PER_R = 0x30644full;       // r < 0x30644f * 2^232
constexpr uint64_t SUBK_LO = (1ull << 30);        // sub<C> adds K[i] < 2^30 to limbs 0..7

constexpr Bnd bnd_normalized(uint32_t v) { return Bnd{LIMB_MASK, v * TOP_PER_R, v, true}; }
constexpr bool mul_ok(const Bnd& b) { return b.lo <= MUL_IN_MAX && b.top <= MUL_IN_MAX && b.v <= MUL_IN_VMAX; }

// Normalisation schedule of a radix-2 DIF network on R registers, decided at
// compile time from the bounds: an operand is normalised (or, when its value
// grows too large, value-reduced) only when an operation would otherwise
// break a contract of fr29.hpp (limbs < 2^32, multiplier operands <=
// MUL_A_LIMB_MAX, sub<C> needs a normalized subtrahend with V < (C-1) r).
// Inputs are normalized with V < 6r; outputs satisfy the multiplier contract.
template <int R>
struct DifSchedule {
  static constexpr int LR = Log2<R>::value;
  static constexpr int NB = LR * R / 2;
  // per butterfly: 0 nothing, 1 normalize, 2 reduce_value
  uint8_t fx[NB > 0 ? NB : 1], fy[NB > 0 ? NB : 1];
  uint32_t c[NB > 0 ? NB : 1];
  uint8_t fout[R];
};

template <int R>
constexpr DifSchedule<R> make_dif_schedule() {
  DifSchedule<R> S{};
  Bnd b[R] = {};
  for (int i = 0; i < R; ++i) b[i] = bnd_normalized(6);
  int idx = 0;
  for (int h = R / 2; h >= 1; h /= 2) {
    for (int j = 0; j < h; ++j) {
      for (int blk = 0; blk < R; blk += 2 * h) {
        const int x = blk + j, y = blk + j + h;
        uint8_t fx = 0, fy = 0;
        Bnd bx = b[x], by = b[y];
        auto apply = [](Bnd& t, uint8_t f) {
          if (f == 2) t = bnd_normalized(4);
          else if (f == 1) t = bnd_normalized(t.v);
        };
        auto fits = [&](const Bnd& X, const Bnd& Y) {
          const uint32_t C = Y.v + 1;
          const uint64_t slo = X.lo + Y.lo, stop = X.top + Y.top;
          const uint64_t dlo = X.lo + SUBK_LO, dtop = X.top + C * TOP_PER_R;
          const uint32_t sv = X.v + Y.v, dv = X.v + C;
          bool ok = Y.norm && slo <= LIMB_MAX32 && stop <= LIMB_MAX32 && sv <= MUL_IN_VMAX && dv <= MUL_IN_VMAX;
          if (j != 0) ok = ok && dlo <= MUL_IN_MAX && dtop <= MUL_IN_MAX;
          else ok = ok && dlo <= LIMB_MAX32 && dtop <= LIMB_MAX32;
          return ok;
        };
        // escalate: normalize y, then x, then reduce both
        if (!by.norm) { fy = 1; apply(by, 1); }
        if (!fits(bx, by) && !bx.norm) { fx = 1; apply(bx, 1); }
        if (!fits(bx, by)) { fx = 2; apply(bx, 2); }
        if (!fits(bx, by)) { fy = 2; apply(by, 2); }
        const uint32_t C = by.v + 1;
        S.fx[idx] = fx;
        S.fy[idx] = fy;
        S.c[idx] = C;
        Bnd s{bx.lo + by.lo, bx.top + by.top, bx.v + by.v, false};
        Bnd d{};
        if (j != 0) d = bnd_normalized(2);
        else d = Bnd{bx.lo + SUBK_LO, bx.top + C * TOP_PER_R, bx.v + C, false};
        b[x] = s;
        b[y] = d;
        ++idx;
      }
    }
  }
  for (int i = 0; i < R; ++i) {
    S.fout[i] = mul_ok(b[i]) ? 0 : (b[i].v > MUL_IN_VMAX ? 2 : 1);
  }
  return S;
}

template <int R>
struct DifSched {
  static constexpr DifSchedule<R> value = make_dif_schedule<R>();
};

template <uint8_t F>
ZK_HD void fix_bound(fe& a) {
  if constexpr (F == 1) normalize(a);
  else if constexpr (F == 2) reduce_value(a);
}

// the sub<C> constants are instantiated per butterfly; C is at most MUL_IN_VMAX + 1
template <int S, int R, int H, bool CHAIN>
ZK_HD void dif_stage(fe* u, const uint4* roots, uint32_t root_log) {
  constexpr int LR = Log2<R>::value;
  constexpr int BASE = S * (R / 2);
  static_for<H>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    fe w;
    if constexpr (j != 0) w = load_limbs48(roots, (uint64_t)(j * (R / (2 * H))) << (root_log - LR));
    static_for<R / (2 * H)>([&](auto bc) {
      constexpr int bi = decltype(bc)::value;
      constexpr int b = bi * 2 * H;
      constexpr int idx = BASE + j * (R / (2 * H)) + bi;
      constexpr uint8_t FX = DifSched<R>::value.fx[idx];
      constexpr uint8_t FY = DifSched<R>::value.fy[idx];
      constexpr uint32_t C = DifSched<R>::value.c[idx];
      fix_bound<FX>(u[b + j]);
      fix_bound<FY>(u[b + j + H]);
      const fe x = u[b + j], y = u[b + j + H];
      u[b + j] = add(x, y);
      if constexpr (j == 0) u[b + j + H] = sub<C>(x, y);
      else u[b + j + H] = mont_mul_u<CHAIN>(sub<C>(x, y), w);
    });
  });
}

// Length-R DFT in registers: natural order in, bit-reversed order out.
// omega_R^e = roots[e << (root_log - log R)]. Inputs normalized with V < 6r;
// outputs satisfy the multiplier's operand contract.
template <int R, bool CHAIN>
ZK_HD void dft_regs(fe* u, const uint4* roots, uint32_t root_log) {
  constexpr int LR = Log2<R>::value;
  static_for<LR>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    dif_stage<s, R, (R >> (s + 1)), CHAIN>(u, roots, root_log);
  });
  static_for<R>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    fix_bound<DifSched<R>::value.fout[i]>(u[i]);
  });
}

}  // namespace zkgpu
