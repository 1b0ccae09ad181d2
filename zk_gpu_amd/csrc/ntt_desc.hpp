// Pass descriptor shared by the host planner and the tile kernels.
//
// A transform of n = 2^k elements runs as P passes with digit sizes
// n_1 ... n_P (n = n_1 * ... * n_P). Pass p applies length-n_p DFTs to the
// current sub-problems; every workgroup ("tile") owns C columns of one
// sub-problem, i.e. C independent length-N DFTs (N = n_p), reads them,
// transforms them through LDS and writes them back. All offsets are in
// elements (32 bytes each). For tile (tlo, thi), column c and position t:
//
//   input  element  in  [tlo*in_lo  + thi*in_hi  + c*in_cs  + t*in_ts ]
//   output element  out [tlo*out_lo + thi*out_hi + c*out_cs + t*out_ts]
//   post-twiddle    wtab[tlo*w_lo   + thi*w_hi   + c*w_cs   + t*w_ts  ]
//
// where t is the input position on load and the output frequency on store.
// pre (load side) is indexed like the input, post (store side) like the
// output; both index the natural order of the whole job.
#pragma once
#include <stdint.h>

namespace zkgpu {

enum : uint32_t {
  PASS_CANON = 1u,    // bring outputs to canonical [0, r)
  PASS_USCALE = 2u,   // multiply outputs by the uniform constant `scale`
};

struct PassDesc {
  const void* in;
  void* out;
  const void* roots;  // omega_Nt^e for e < Nt (Nt = 2^root_log), Montgomery
  const void* wtab;   // inter-pass twiddles or nullptr
  const void* pre;    // load-side per-element scale or nullptr
  const void* post;   // store-side per-element scale or nullptr
  uint32_t scale[12];  // uniform store-side scale, 9 radix-2^29 limbs in the 2^261 domain (PASS_USCALE)
  uint32_t log_n_tile;
  uint32_t log_c;
  uint32_t root_log;
  uint32_t log_n_lo;
  uint32_t num_tiles;
  uint32_t flags;
  uint32_t in_lo, in_hi, in_ts, in_cs;
  uint32_t out_lo, out_hi, out_ts, out_cs;
  uint32_t w_lo, w_hi, w_ts, w_cs;
};

// Largest tile (elements): N * C <= 2^TILE_LOG_MAX. 4096 x 32 B = 128 KiB of LDS.
constexpr uint32_t TILE_LOG_MAX = 12;
constexpr uint32_t ELEMS_PER_THREAD_LOG = 4;  // radix-16 rounds, 16 elements per thread

}  // namespace zkgpu
