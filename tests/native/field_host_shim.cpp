// Host build of the library's field code for CPU tests (tests/test_field_host.py).
#include <string.h>

#include "../../zk_gpu_amd/csrc/fr29.hpp"
#include "../../zk_gpu_amd/csrc/host_field.hpp"

using namespace zkgpu;

static fe ld(const uint32_t* p) { fe a; memcpy(a.l, p, sizeof(a.l)); return a; }
static void st(uint32_t* p, const fe& a) { memcpy(p, a.l, sizeof(a.l)); }

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
