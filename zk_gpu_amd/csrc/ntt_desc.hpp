// Pass descriptor shared by the host planner and the tile kernels.
//
// A transform of n = 2^k elements runs as P passes with digit sizes
// n_1 ... n_P (n = n_1 * ... * n_P). Pass p applies length-n_p DFTs to the
// current sub-problems; every workgroup ("tile") owns C columns of one
// sub-problem, i.e. C independent length-N DFTs (N = n_p), reads them,
// transforms them through LDS and writes them back. All offsets are in
// elements. For tile (tlo, thi), column c and position t:
//
//   input  element  in  [tlo*in_lo  + thi*in_hi  + c*in_cs  + t*in_ts ]
//   output element  out [tlo*out_lo + thi*out_hi + c*out_cs + t*out_ts]
//
// where t is the input position on load and the output frequency on store.
// `pre` is multiplied into every loaded element, indexed by its input offset
// masked with pre_mask: the coset powers c^i in the first pass, the previous
// pass's inter-pass twiddles (laid out like that pass's output) otherwise.
// `post` (coset-inverse powers) is indexed by the output offset.
//
// Element formats: packed (4 little-endian u64 words, value < 2^256, the
// caller's layout) or planar (the workspace between passes: the nine 29-bit
// limbs as a uint4 plane of limbs 0-3, a uint4 plane of limbs 4-7 and a u32
// plane of limb 8, each plane `*_plane` elements long).
#pragma once
#include <stdint.h>

namespace zkgpu {

enum : uint32_t {
  PASS_CANON = 1u,       // bring outputs to canonical [0, r)
  PASS_USCALE = 2u,      // multiply outputs by the uniform constant `scale`
  PASS_IN_PLANAR = 4u,   // input in planar limb form
  PASS_OUT_PLANAR = 8u,  // output in planar limb form
};

struct PassDesc {
  const void* in;
  void* out;
  const void* roots;  // omega_Nt^e for e < Nt (Nt = 2^root_log), 48-byte limb entries, 2^261 domain
  const void* pre;    // load-side per-element factor or nullptr (packed, 2^261 domain)
  const void* post;   // store-side per-element factor or nullptr (packed, 2^261 domain)
  uint32_t scale[12];  // uniform store-side scale, 9 radix-2^29 limbs in the 2^261 domain (PASS_USCALE)
  uint32_t log_n_tile;
  uint32_t log_c;
  uint32_t root_log;
  uint32_t log_n_lo;
  uint32_t num_tiles;
  uint32_t flags;
  uint32_t pre_mask;
  uint32_t in_plane, out_plane;
  uint32_t in_lo, in_hi, in_ts, in_cs;
  uint32_t out_lo, out_hi, out_ts, out_cs;
};

// Largest tile (elements): N * C <= 2^TILE_LOG_MAX; 4096 x 20 B = 80 KiB of LDS.
constexpr uint32_t TILE_LOG_MAX = 12;

}  // namespace zkgpu
