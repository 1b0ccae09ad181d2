// Host build of the library's field code for CPU tests (tests/test_field_host.py).
#include <string.h>

#include <array>
#include <utility>

#include "../../zk_gpu_amd/csrc/dif.hpp"
#include "../../zk_gpu_amd/csrc/fr29.hpp"
#include "../../zk_gpu_amd/csrc/host_field.hpp"

using namespace zkgpu;

static fe ld(const uint32_t* p) { fe a; memcpy(a.l, p, sizeof(a.l)); return a; }
static void st(uint32_t* p, const fe& a) { memcpy(p, a.l, sizeof(a.l)); }

// sub<C> and its constant for every C below kMaxC (the DIF schedule
// instantiates C up to MUL_IN_VMAX + 1)
constexpr uint32_t kMaxC = 200;
using SubFn = void (*)(const uint32_t*, const uint32_t*, uint32_t*);
using KFn = uint32_t (*)(int);
template <uint32_t C>
void sub_c(const uint32_t* a, const uint32_t* b, uint32_t* o) { st(o, sub<C>(ld(a), ld(b))); }
template <uint32_t C>
uint32_t k_c(int i) { return SubConst<C>::k(i); }
template <uint32_t... Cs>
constexpr std::array<SubFn, sizeof...(Cs)> sub_table(std::integer_sequence<uint32_t, Cs...>) { return {{&sub_c<Cs>...}}; }
template <uint32_t... Cs>
constexpr std::array<KFn, sizeof...(Cs)> k_table(std::integer_sequence<uint32_t, Cs...>) { return {{&k_c<Cs>...}}; }
static const auto SUB_TABLE = sub_table(std::make_integer_sequence<uint32_t, kMaxC>{});
static const auto K_TABLE = k_table(std::make_integer_sequence<uint32_t, kMaxC>{});

// The register DFT of the pass kernels, built for the host.
template <int R, bool CHAIN>
static void dft_host(const uint32_t* in, const uint32_t* roots, uint32_t root_log, uint32_t* out) {
  alignas(16) uint32_t tab[12 * 64];
  memcpy(tab, roots, sizeof(uint32_t) * 12 * (1u << root_log));
  fe u[R];
  for (int i = 0; i < R; ++i) u[i] = ld(in + 9 * i);
  dft_regs<R, CHAIN>(u, (const uint4*)tab, root_log);
  for (int i = 0; i < R; ++i) st(out + 9 * i, u[i]);
}
template <int R>
static void schedule_host(uint32_t* out) {
  // per butterfly: fx, fy, C; then fout per register
  constexpr auto S = DifSched<R>::value;
  constexpr int NB = DifSchedule<R>::NB;
  int o = 0;
  for (int b = 0; b < NB; ++b) {
    out[o++] = S.fx[b];
    out[o++] = S.fy[b];
    out[o++] = S.c[b];
  }
  for (int i = 0; i < R; ++i) out[o++] = S.fout[i];
}

extern "C" {
void shim_mont_mul(const uint32_t* a, const uint32_t* b, uint32_t* o) { st(o, mont_mul(ld(a), ld(b))); }
void shim_add(const uint32_t* a, const uint32_t* b, uint32_t* o) { st(o, add(ld(a), ld(b))); }
void shim_sub(uint32_t c, const uint32_t* a, const uint32_t* b, uint32_t* o) {
  switch (c) {
    case 5: st(o, sub<5>(ld(a), ld(b))); break;
    case 10: st(o, sub<10>(ld(a), ld(b))); break;
    case 20: st(o, sub<20>(ld(a), ld(b))); break;
    case 40: st(o, sub<40>(ld(a), ld(b))); break;
    default: memset(o, 0xff, 36);
  }
}
int shim_sub_any(uint32_t c, const uint32_t* a, const uint32_t* b, uint32_t* o) {
  if (c >= kMaxC) return -1;
  SUB_TABLE[c](a, b, o);
  return 0;
}
int shim_subconst_any(uint32_t c, uint32_t* out) {
  if (c >= kMaxC) return -1;
  for (int i = 0; i < 9; ++i) out[i] = K_TABLE[c](i);
  return 0;
}
int shim_dft(int R, int chain, const uint32_t* in, const uint32_t* roots, uint32_t root_log, uint32_t* out) {
  if (root_log > 6) return -1;
  switch (R * 2 + (chain ? 1 : 0)) {
    case 4: dft_host<2, false>(in, roots, root_log, out); return 0;
    case 5: dft_host<2, true>(in, roots, root_log, out); return 0;
    case 8: dft_host<4, false>(in, roots, root_log, out); return 0;
    case 9: dft_host<4, true>(in, roots, root_log, out); return 0;
    case 16: dft_host<8, false>(in, roots, root_log, out); return 0;
    case 17: dft_host<8, true>(in, roots, root_log, out); return 0;
    default: return -1;
  }
}
// the schedule of the radix-R network; returns the number of words written
int shim_dif_schedule(int R, uint32_t* out) {
  switch (R) {
    case 2: schedule_host<2>(out); return 3 * DifSchedule<2>::NB + 2;
    case 4: schedule_host<4>(out); return 3 * DifSchedule<4>::NB + 4;
    case 8: schedule_host<8>(out); return 3 * DifSchedule<8>::NB + 8;
    default: return -1;
  }
}
void shim_mul_contract(uint32_t* out) { out[0] = (uint32_t)MUL_IN_MAX; out[1] = MUL_IN_VMAX; }
void shim_normalize(uint32_t* a) { fe v = ld(a); normalize(v); st(a, v); }
void shim_reduce_value(uint32_t* a) { fe v = ld(a); reduce_value(v); st(a, v); }
void shim_pack(const uint32_t* a, uint32_t* w) { pack_words(w, ld(a)); }
void shim_unpack(const uint32_t* w, uint32_t* a) { st(a, unpack_words(w)); }
void shim_to_canonical(const uint32_t* a, uint32_t* w) { to_canonical_words(w, ld(a)); }
void shim_to_limbs261(const uint64_t* x, uint32_t* out) { host::to_limbs261(out, host::from_ptr(x)); }
void shim_host_mul(const uint64_t* a, const uint64_t* b, uint64_t* o) {
  host::Fr r = host::mul(host::from_ptr(a), host::from_ptr(b)); memcpy(o, r.w, 32);
}
void shim_host_inv(const uint64_t* a, uint64_t* o) { host::Fr r = host::inv(host::from_ptr(a)); memcpy(o, r.w, 32); }
void shim_root(uint32_t k, uint64_t* o) { host::Fr r = host::root_of_unity(k); memcpy(o, r.w, 32); }
void shim_consts(uint32_t* out) {
  // TO261, ONE261, ENCODE256, QMAGIC, MUL_A_LIMB_MAX
  memcpy(out, TO261, 36); memcpy(out + 9, ONE261, 36); memcpy(out + 18, ENCODE256, 36);
  out[27] = QMAGIC; out[28] = MUL_A_LIMB_MAX;
}
void shim_subconst(uint32_t c, uint32_t* out) {
  for (int i = 0; i < 9; ++i) {
    switch (c) {
      case 5: out[i] = SubConst<5>::k(i); break;
      case 10: out[i] = SubConst<10>::k(i); break;
      case 20: out[i] = SubConst<20>::k(i); break;
      case 40: out[i] = SubConst<40>::k(i); break;
    }
  }
}
}
