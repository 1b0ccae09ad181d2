// Generated by gen_mont_cols.py; do not edit.
//
// mont_mul_cols<M, SB>(a, b) = a * b * 2^-261 mod M::p (not reduced),
// the product scanning of zkgpu::mont_mul with every column one chain of
// v_mad_u64_u32 in a single asm statement (carry-out to VCC, unused).
// M provides p(i), the radix-2^29 limbs of the modulus, and ninv =
// -p^-1 mod 2^29. SB: b is the same in every lane of the wave and its
// limbs are taken from SGPRs (a value held in VGPRs is read from the
// first active lane, so SB is only for wave-uniform operands). Same
// contracts as mont_mul. Device code only.
#pragma once
#include <stdint.h>

namespace zkgpu {

template <class M, bool SB = false>
__device__ __forceinline__ fe mont_mul_cols(const fe& a, const fe& b) {
  uint32_t m[NLIMBS];
  fe o;
  uint64_t acc = (uint64_t)a.l[0] * b.l[0];
  m[0] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[0] * M::p(0);
  acc >>= LIMB_BITS;
  // column 1: 3 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[1]), "v"(a.l[1]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[1]), "v"(a.l[1]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(1))
        : "vcc");
  }
  m[1] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[1] * M::p(0);
  acc >>= LIMB_BITS;
  // column 2: 5 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[2]), "v"(a.l[1]), "v"(b.l[1]), "v"(a.l[2]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(2)), "v"(m[1]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[2]), "v"(a.l[1]), "s"(b.l[1]), "v"(a.l[2]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(2)), "v"(m[1]), "s"(M::p(1))
        : "vcc");
  }
  m[2] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[2] * M::p(0);
  acc >>= LIMB_BITS;
  // column 3: 7 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[3]), "v"(a.l[1]), "v"(b.l[2]), "v"(a.l[2]), "v"(b.l[1]), "v"(a.l[3]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(3)), "v"(m[1]), "s"(M::p(2)), "v"(m[2]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[3]), "v"(a.l[1]), "s"(b.l[2]), "v"(a.l[2]), "s"(b.l[1]), "v"(a.l[3]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(3)), "v"(m[1]), "s"(M::p(2)), "v"(m[2]), "s"(M::p(1))
        : "vcc");
  }
  m[3] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[3] * M::p(0);
  acc >>= LIMB_BITS;
  // column 4: 9 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[4]), "v"(a.l[1]), "v"(b.l[3]), "v"(a.l[2]), "v"(b.l[2]), "v"(a.l[3]), "v"(b.l[1]), "v"(a.l[4]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(4)), "v"(m[1]), "s"(M::p(3)), "v"(m[2]), "s"(M::p(2)), "v"(m[3]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[4]), "v"(a.l[1]), "s"(b.l[3]), "v"(a.l[2]), "s"(b.l[2]), "v"(a.l[3]), "s"(b.l[1]), "v"(a.l[4]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(4)), "v"(m[1]), "s"(M::p(3)), "v"(m[2]), "s"(M::p(2)), "v"(m[3]), "s"(M::p(1))
        : "vcc");
  }
  m[4] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[4] * M::p(0);
  acc >>= LIMB_BITS;
  // column 5: 11 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[5]), "v"(a.l[1]), "v"(b.l[4]), "v"(a.l[2]), "v"(b.l[3]), "v"(a.l[3]), "v"(b.l[2]), "v"(a.l[4]), "v"(b.l[1]), "v"(a.l[5]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(5)), "v"(m[1]), "s"(M::p(4)), "v"(m[2]), "s"(M::p(3)), "v"(m[3]), "s"(M::p(2)), "v"(m[4]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[5]), "v"(a.l[1]), "s"(b.l[4]), "v"(a.l[2]), "s"(b.l[3]), "v"(a.l[3]), "s"(b.l[2]), "v"(a.l[4]), "s"(b.l[1]), "v"(a.l[5]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(5)), "v"(m[1]), "s"(M::p(4)), "v"(m[2]), "s"(M::p(3)), "v"(m[3]), "s"(M::p(2)), "v"(m[4]), "s"(M::p(1))
        : "vcc");
  }
  m[5] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[5] * M::p(0);
  acc >>= LIMB_BITS;
  // column 6: 13 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[6]), "v"(a.l[1]), "v"(b.l[5]), "v"(a.l[2]), "v"(b.l[4]), "v"(a.l[3]), "v"(b.l[3]), "v"(a.l[4]), "v"(b.l[2]), "v"(a.l[5]), "v"(b.l[1]), "v"(a.l[6]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(6)), "v"(m[1]), "s"(M::p(5)), "v"(m[2]), "s"(M::p(4)), "v"(m[3]), "s"(M::p(3)), "v"(m[4]), "s"(M::p(2)), "v"(m[5]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[6]), "v"(a.l[1]), "s"(b.l[5]), "v"(a.l[2]), "s"(b.l[4]), "v"(a.l[3]), "s"(b.l[3]), "v"(a.l[4]), "s"(b.l[2]), "v"(a.l[5]), "s"(b.l[1]), "v"(a.l[6]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(6)), "v"(m[1]), "s"(M::p(5)), "v"(m[2]), "s"(M::p(4)), "v"(m[3]), "s"(M::p(3)), "v"(m[4]), "s"(M::p(2)), "v"(m[5]), "s"(M::p(1))
        : "vcc");
  }
  m[6] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[6] * M::p(0);
  acc >>= LIMB_BITS;
  // column 7: 15 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[7]), "v"(a.l[1]), "v"(b.l[6]), "v"(a.l[2]), "v"(b.l[5]), "v"(a.l[3]), "v"(b.l[4]), "v"(a.l[4]), "v"(b.l[3]), "v"(a.l[5]), "v"(b.l[2]), "v"(a.l[6]), "v"(b.l[1]), "v"(a.l[7]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(7)), "v"(m[1]), "s"(M::p(6)), "v"(m[2]), "s"(M::p(5)), "v"(m[3]), "s"(M::p(4)), "v"(m[4]), "s"(M::p(3)), "v"(m[5]), "s"(M::p(2)), "v"(m[6]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[7]), "v"(a.l[1]), "s"(b.l[6]), "v"(a.l[2]), "s"(b.l[5]), "v"(a.l[3]), "s"(b.l[4]), "v"(a.l[4]), "s"(b.l[3]), "v"(a.l[5]), "s"(b.l[2]), "v"(a.l[6]), "s"(b.l[1]), "v"(a.l[7]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(7)), "v"(m[1]), "s"(M::p(6)), "v"(m[2]), "s"(M::p(5)), "v"(m[3]), "s"(M::p(4)), "v"(m[4]), "s"(M::p(3)), "v"(m[5]), "s"(M::p(2)), "v"(m[6]), "s"(M::p(1))
        : "vcc");
  }
  m[7] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[7] * M::p(0);
  acc >>= LIMB_BITS;
  // column 8: 17 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %31, %32, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %33, %34, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "v"(b.l[8]), "v"(a.l[1]), "v"(b.l[7]), "v"(a.l[2]), "v"(b.l[6]), "v"(a.l[3]), "v"(b.l[5]), "v"(a.l[4]), "v"(b.l[4]), "v"(a.l[5]), "v"(b.l[3]), "v"(a.l[6]), "v"(b.l[2]), "v"(a.l[7]), "v"(b.l[1]), "v"(a.l[8]), "v"(b.l[0]), "v"(m[0]), "s"(M::p(8)), "v"(m[1]), "s"(M::p(7)), "v"(m[2]), "s"(M::p(6)), "v"(m[3]), "s"(M::p(5)), "v"(m[4]), "s"(M::p(4)), "v"(m[5]), "s"(M::p(3)), "v"(m[6]), "s"(M::p(2)), "v"(m[7]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %31, %32, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %33, %34, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[0]), "s"(b.l[8]), "v"(a.l[1]), "s"(b.l[7]), "v"(a.l[2]), "s"(b.l[6]), "v"(a.l[3]), "s"(b.l[5]), "v"(a.l[4]), "s"(b.l[4]), "v"(a.l[5]), "s"(b.l[3]), "v"(a.l[6]), "s"(b.l[2]), "v"(a.l[7]), "s"(b.l[1]), "v"(a.l[8]), "s"(b.l[0]), "v"(m[0]), "s"(M::p(8)), "v"(m[1]), "s"(M::p(7)), "v"(m[2]), "s"(M::p(6)), "v"(m[3]), "s"(M::p(5)), "v"(m[4]), "s"(M::p(4)), "v"(m[5]), "s"(M::p(3)), "v"(m[6]), "s"(M::p(2)), "v"(m[7]), "s"(M::p(1))
        : "vcc");
  }
  m[8] = ((uint32_t)acc * M::ninv) & LIMB_MASK;
  acc += (uint64_t)m[8] * M::p(0);
  acc >>= LIMB_BITS;
  // column 9: 16 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %31, %32, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[1]), "v"(b.l[8]), "v"(a.l[2]), "v"(b.l[7]), "v"(a.l[3]), "v"(b.l[6]), "v"(a.l[4]), "v"(b.l[5]), "v"(a.l[5]), "v"(b.l[4]), "v"(a.l[6]), "v"(b.l[3]), "v"(a.l[7]), "v"(b.l[2]), "v"(a.l[8]), "v"(b.l[1]), "v"(m[1]), "s"(M::p(8)), "v"(m[2]), "s"(M::p(7)), "v"(m[3]), "s"(M::p(6)), "v"(m[4]), "s"(M::p(5)), "v"(m[5]), "s"(M::p(4)), "v"(m[6]), "s"(M::p(3)), "v"(m[7]), "s"(M::p(2)), "v"(m[8]), "s"(M::p(1))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %29, %30, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %31, %32, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[1]), "s"(b.l[8]), "v"(a.l[2]), "s"(b.l[7]), "v"(a.l[3]), "s"(b.l[6]), "v"(a.l[4]), "s"(b.l[5]), "v"(a.l[5]), "s"(b.l[4]), "v"(a.l[6]), "s"(b.l[3]), "v"(a.l[7]), "s"(b.l[2]), "v"(a.l[8]), "s"(b.l[1]), "v"(m[1]), "s"(M::p(8)), "v"(m[2]), "s"(M::p(7)), "v"(m[3]), "s"(M::p(6)), "v"(m[4]), "s"(M::p(5)), "v"(m[5]), "s"(M::p(4)), "v"(m[6]), "s"(M::p(3)), "v"(m[7]), "s"(M::p(2)), "v"(m[8]), "s"(M::p(1))
        : "vcc");
  }
  o.l[0] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 10: 14 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[2]), "v"(b.l[8]), "v"(a.l[3]), "v"(b.l[7]), "v"(a.l[4]), "v"(b.l[6]), "v"(a.l[5]), "v"(b.l[5]), "v"(a.l[6]), "v"(b.l[4]), "v"(a.l[7]), "v"(b.l[3]), "v"(a.l[8]), "v"(b.l[2]), "v"(m[2]), "s"(M::p(8)), "v"(m[3]), "s"(M::p(7)), "v"(m[4]), "s"(M::p(6)), "v"(m[5]), "s"(M::p(5)), "v"(m[6]), "s"(M::p(4)), "v"(m[7]), "s"(M::p(3)), "v"(m[8]), "s"(M::p(2))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %25, %26, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %27, %28, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[2]), "s"(b.l[8]), "v"(a.l[3]), "s"(b.l[7]), "v"(a.l[4]), "s"(b.l[6]), "v"(a.l[5]), "s"(b.l[5]), "v"(a.l[6]), "s"(b.l[4]), "v"(a.l[7]), "s"(b.l[3]), "v"(a.l[8]), "s"(b.l[2]), "v"(m[2]), "s"(M::p(8)), "v"(m[3]), "s"(M::p(7)), "v"(m[4]), "s"(M::p(6)), "v"(m[5]), "s"(M::p(5)), "v"(m[6]), "s"(M::p(4)), "v"(m[7]), "s"(M::p(3)), "v"(m[8]), "s"(M::p(2))
        : "vcc");
  }
  o.l[1] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 11: 12 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[3]), "v"(b.l[8]), "v"(a.l[4]), "v"(b.l[7]), "v"(a.l[5]), "v"(b.l[6]), "v"(a.l[6]), "v"(b.l[5]), "v"(a.l[7]), "v"(b.l[4]), "v"(a.l[8]), "v"(b.l[3]), "v"(m[3]), "s"(M::p(8)), "v"(m[4]), "s"(M::p(7)), "v"(m[5]), "s"(M::p(6)), "v"(m[6]), "s"(M::p(5)), "v"(m[7]), "s"(M::p(4)), "v"(m[8]), "s"(M::p(3))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %21, %22, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %23, %24, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[3]), "s"(b.l[8]), "v"(a.l[4]), "s"(b.l[7]), "v"(a.l[5]), "s"(b.l[6]), "v"(a.l[6]), "s"(b.l[5]), "v"(a.l[7]), "s"(b.l[4]), "v"(a.l[8]), "s"(b.l[3]), "v"(m[3]), "s"(M::p(8)), "v"(m[4]), "s"(M::p(7)), "v"(m[5]), "s"(M::p(6)), "v"(m[6]), "s"(M::p(5)), "v"(m[7]), "s"(M::p(4)), "v"(m[8]), "s"(M::p(3))
        : "vcc");
  }
  o.l[2] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 12: 10 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[4]), "v"(b.l[8]), "v"(a.l[5]), "v"(b.l[7]), "v"(a.l[6]), "v"(b.l[6]), "v"(a.l[7]), "v"(b.l[5]), "v"(a.l[8]), "v"(b.l[4]), "v"(m[4]), "s"(M::p(8)), "v"(m[5]), "s"(M::p(7)), "v"(m[6]), "s"(M::p(6)), "v"(m[7]), "s"(M::p(5)), "v"(m[8]), "s"(M::p(4))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %17, %18, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %19, %20, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[4]), "s"(b.l[8]), "v"(a.l[5]), "s"(b.l[7]), "v"(a.l[6]), "s"(b.l[6]), "v"(a.l[7]), "s"(b.l[5]), "v"(a.l[8]), "s"(b.l[4]), "v"(m[4]), "s"(M::p(8)), "v"(m[5]), "s"(M::p(7)), "v"(m[6]), "s"(M::p(6)), "v"(m[7]), "s"(M::p(5)), "v"(m[8]), "s"(M::p(4))
        : "vcc");
  }
  o.l[3] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 13: 8 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[5]), "v"(b.l[8]), "v"(a.l[6]), "v"(b.l[7]), "v"(a.l[7]), "v"(b.l[6]), "v"(a.l[8]), "v"(b.l[5]), "v"(m[5]), "s"(M::p(8)), "v"(m[6]), "s"(M::p(7)), "v"(m[7]), "s"(M::p(6)), "v"(m[8]), "s"(M::p(5))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %15, %16, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[5]), "s"(b.l[8]), "v"(a.l[6]), "s"(b.l[7]), "v"(a.l[7]), "s"(b.l[6]), "v"(a.l[8]), "s"(b.l[5]), "v"(m[5]), "s"(M::p(8)), "v"(m[6]), "s"(M::p(7)), "v"(m[7]), "s"(M::p(6)), "v"(m[8]), "s"(M::p(5))
        : "vcc");
  }
  o.l[4] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 14: 6 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[6]), "v"(b.l[8]), "v"(a.l[7]), "v"(b.l[7]), "v"(a.l[8]), "v"(b.l[6]), "v"(m[6]), "s"(M::p(8)), "v"(m[7]), "s"(M::p(7)), "v"(m[8]), "s"(M::p(6))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %11, %12, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[6]), "s"(b.l[8]), "v"(a.l[7]), "s"(b.l[7]), "v"(a.l[8]), "s"(b.l[6]), "v"(m[6]), "s"(M::p(8)), "v"(m[7]), "s"(M::p(7)), "v"(m[8]), "s"(M::p(6))
        : "vcc");
  }
  o.l[5] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 15: 4 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[7]), "v"(b.l[8]), "v"(a.l[8]), "v"(b.l[7]), "v"(m[7]), "s"(M::p(8)), "v"(m[8]), "s"(M::p(7))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %7, %8, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[7]), "s"(b.l[8]), "v"(a.l[8]), "s"(b.l[7]), "v"(m[7]), "s"(M::p(8)), "v"(m[8]), "s"(M::p(7))
        : "vcc");
  }
  o.l[6] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  // column 16: 2 multiply-adds
  if constexpr (!SB) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[8]), "v"(b.l[8]), "v"(m[8]), "s"(M::p(8))
        : "vcc");
  } else {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0\n\t"
        "v_mad_u64_u32 %0, vcc, %3, %4, %0\n\t"
        : "+v"(acc)
        : "v"(a.l[8]), "s"(b.l[8]), "v"(m[8]), "s"(M::p(8))
        : "vcc");
  }
  o.l[7] = (uint32_t)acc & LIMB_MASK;
  acc >>= LIMB_BITS;
  o.l[NLIMBS - 1] = (uint32_t)acc;
  return o;
}

}  // namespace zkgpu
