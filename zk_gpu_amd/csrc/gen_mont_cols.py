"""Writes mont_cols.hpp: the device Montgomery product with each column's
multiply-adds as one inline-asm chain of v_mad_u64_u32.

    python zk_gpu_amd/csrc/gen_mont_cols.py        # rewrite the header
    python zk_gpu_amd/csrc/gen_mont_cols.py --check  # exit 1 if it is stale

Why asm: a column accumulated as one dependent chain (the carry first, then
every partial product) needs no partial-sum merges, but the compiler
reassociates such a chain into pieces that start from zero. An empty asm
statement on the accumulator keeps the chain, and the compiler then puts an
s_nop before nearly every multiply-add (it must assume the empty asm could
be an instruction whose result needs a wait state): 143 nops per product.
With a whole column in one asm statement the wait state is needed only once
per column.
"""
import os
import sys

NL = 9
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "mont_cols.hpp")


def column_terms(k):
    """(kind, i, j) of column k: a[i]*b[j] with i + j = k, then m[i]*p[j] with
    i + j = k, i < k, j >= 1 (the m[k]*p[0] term follows the digit)."""
    ab = [("ab", i, k - i) for i in range(NL) if 0 <= k - i < NL]
    mp = [("mp", i, k - i) for i in range(NL) if i < k and 1 <= k - i < NL]
    return ab + mp


def render():
    out = [
        "// Generated by gen_mont_cols.py; do not edit.",
        "//",
        "// mont_mul_cols<M, SB>(a, b) = a * b * 2^-261 mod M::p (not reduced),",
        "// the product scanning of zkgpu::mont_mul with every column one chain of",
        "// v_mad_u64_u32 in a single asm statement (carry-out to VCC, unused).",
        "// M provides p(i), the radix-2^29 limbs of the modulus, and ninv =",
        "// -p^-1 mod 2^29. SB: b is the same in every lane of the wave and its",
        "// limbs are taken from SGPRs (a value held in VGPRs is read from the",
        "// first active lane, so SB is only for wave-uniform operands). Same",
        "// contracts as mont_mul. Device code only.",
        "#pragma once",
        "#include <stdint.h>",
        "",
        "namespace zkgpu {",
        "",
        "template <class M, bool SB = false>",
        "__device__ __forceinline__ fe mont_mul_cols(const fe& a, const fe& b) {",
        "  uint32_t m[NLIMBS];",
        "  fe o;",
        "  uint64_t acc = (uint64_t)a.l[0] * b.l[0];",
        "  m[0] = ((uint32_t)acc * M::ninv) & LIMB_MASK;",
        "  acc += (uint64_t)m[0] * M::p(0);",
        "  acc >>= LIMB_BITS;",
    ]
    for k in range(1, 2 * NL - 1):
        terms = column_terms(k)
        out.append(f"  // column {k}: {len(terms)} multiply-adds")
        for sb in (False, True):
            lines, ins = [], []
            n = 1
            for kind, i, j in terms:
                lines.append(f'"v_mad_u64_u32 %0, vcc, %{n}, %{n + 1}, %0\\n\\t"')
                if kind == "ab":
                    ins += [f'"v"(a.l[{i}])', f'"{"s" if sb else "v"}"(b.l[{j}])']
                else:
                    ins += [f'"v"(m[{i}])', f'"s"(M::p({j}))']
                n += 2
            out.append(f"  {'} else {' if sb else 'if constexpr (!SB) {'}")
            out.append("    asm(" + "\n        ".join(lines))
            out.append('        : "+v"(acc)')
            out.append("        : " + ", ".join(ins))
            out.append('        : "vcc");')
        out.append("  }")
        if k < NL:
            out.append(f"  m[{k}] = ((uint32_t)acc * M::ninv) & LIMB_MASK;")
            out.append(f"  acc += (uint64_t)m[{k}] * M::p(0);")
        else:
            out.append(f"  o.l[{k - NL}] = (uint32_t)acc & LIMB_MASK;")
        out.append("  acc >>= LIMB_BITS;")
    out += [
        "  o.l[NLIMBS - 1] = (uint32_t)acc;",
        "  return o;",
        "}",
        "",
        "}  // namespace zkgpu",
        "",
    ]
    return "\n".join(out)


def main():
    text = render()
    if "--check" in sys.argv:
        with open(OUT) as f:
            if f.read() != text:
                print(f"{OUT} is stale: run {__file__}")
                sys.exit(1)
        return
    with open(OUT, "w") as f:
        f.write(text)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
