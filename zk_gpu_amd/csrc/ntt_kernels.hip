// HIP kernels of the BN256 Fr NTT for gfx950 (MI355X).
//
// ntt_tile: one workgroup owns C columns x N positions of one pass (see
// ntt_desc.hpp). Each thread keeps 8 elements in 29-bit-limb form
// (8 x 9 VGPRs) and the tile runs Stockham (self-sorting) rounds of radix
// 8 (the last round takes the remaining radix 2/4): round 1 reads HBM,
// middle rounds exchange through LDS (dword planes, conflict-free for
// consecutive positions, limbs in two phases), the last round writes HBM.
// Every pass therefore reads and writes each element once. Inside a round a
// thread multiplies its inputs by the round twiddles omega_{pR}^{jk}, runs a
// radix-2 decimation-in-frequency network over its R registers and scatters
// the (bit-reversal relabelled) outputs.
//
// Pass-level fusions: the coset pre-scale (a[i] *= c^i, reference
// cuda/fft.cu:223-232) on the first load; the inter-pass twiddle, the
// inverse's divisor (reference cuda/fft.cu:153-155) and the coset-inverse
// post-scale on the last store, together with the final canonicalisation.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <type_traits>

#include "dif.hpp"
#include "fr29.hpp"
#include "ntt_desc.hpp"

namespace zkgpu {

// ---------------------------------------------------------------- memory

// packed element: 8 x u32 (value < 2^256), 32 bytes
__device__ __forceinline__ fe load_packed(const void* base, uint64_t idx) {
  const uint4* p = (const uint4*)base + 2 * idx;
  const uint4 lo = p[0], hi = p[1];
  const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  return unpack_words(w);
}
__device__ __forceinline__ void store_words(void* base, uint64_t idx, const uint32_t w[8]) {
  uint4* p = (uint4*)base + 2 * idx;
  p[0] = make_uint4(w[0], w[1], w[2], w[3]);
  p[1] = make_uint4(w[4], w[5], w[6], w[7]);
}
__device__ __forceinline__ void store_limbs48(void* base, uint64_t idx, const fe& v) {
  uint4* p = (uint4*)base + 3 * idx;
  p[0] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  p[1] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
  p[2] = make_uint4(v.l[8], 0u, 0u, 0u);
}
// Planar limb form (workspace between passes and LDS): limbs 0-3 in a uint4
// plane, limbs 4-7 in a second uint4 plane, limb 8 in a u32 plane.
__device__ __forceinline__ fe load_planar(const void* base, uint32_t plane, uint32_t idx) {
  const uint4* a = (const uint4*)base;
  const uint4* b = a + plane;
  const uint32_t* c = (const uint32_t*)(b + plane);
  const uint4 x = a[idx], y = b[idx];
  fe r;
  r.l[0] = x.x; r.l[1] = x.y; r.l[2] = x.z; r.l[3] = x.w;
  r.l[4] = y.x; r.l[5] = y.y; r.l[6] = y.z; r.l[7] = y.w;
  r.l[8] = c[idx];
  return r;
}
__device__ __forceinline__ void store_planar(void* base, uint32_t plane, uint32_t idx, const fe& v) {
  uint4* a = (uint4*)base;
  uint4* b = a + plane;
  uint32_t* c = (uint32_t*)(b + plane);
  a[idx] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  b[idx] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
  c[idx] = v.l[8];
}

// How the rounds of a tile exchange data through LDS.
//   two_phase: limbs 0-3 (a uint4 plane) and limb 4 (a u32 plane), then
//              limbs 5-8 in the same uint4 plane, each phase closed by a
//              barrier: 20 B per element (a 2048-element tile takes 40 KiB
//              instead of 72 KiB, so up to four share a CU), for two more
//              barriers per exchange. Otherwise all nine limbs at once in the
//              planes of store_planar (36 B per element).
//   chain:     the multiplier's single-chain column form (mont_mul<true>,
//              one asm chain of multiply-adds per column).
//   le_max:    log2 of the elements each thread keeps (radix of the rounds).
//   swizzle:   LDS element index with bits 1-3 flipped by bits 4-6 (lds_elem)
// The round twiddles are loaded after the barriers: loading them ahead,
// across the barrier, spilled or ran slower in every configuration.
template <bool TWO_PHASE, bool CHAIN, int LE_MAX = 3, bool SWIZZLE = false>
struct Xchg {
  static constexpr bool two_phase = TWO_PHASE;
  static constexpr bool chain = CHAIN;
  static constexpr int le_max = LE_MAX;
  static constexpr bool swizzle = SWIZZLE;
  static constexpr uint32_t words = TWO_PHASE ? 5u : (uint32_t)NLIMBS;  // LDS words per tile element
};

// PH 0: limbs 0-4, PH 1: limbs 5-8 (the two phases), PH 2: all nine limbs
template <int PH>
__device__ __forceinline__ void lds_put(uint32_t* lds, uint32_t plane, uint32_t idx, const fe& v) {
  uint4* a = (uint4*)lds;
  if constexpr (PH == 2) {
    store_planar(lds, plane, idx, v);
  } else if constexpr (PH == 0) {
    a[idx] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
    ((uint32_t*)(a + plane))[idx] = v.l[4];
  } else {
    a[idx] = make_uint4(v.l[5], v.l[6], v.l[7], v.l[8]);
  }
}
template <int PH>
__device__ __forceinline__ void lds_get(const uint32_t* lds, uint32_t plane, uint32_t idx, fe& v) {
  const uint4* a = (const uint4*)lds;
  if constexpr (PH == 2) {
    v = load_planar(lds, plane, idx);
    return;
  }
  const uint4 x = a[idx];
  if constexpr (PH == 0) {
    v.l[0] = x.x; v.l[1] = x.y; v.l[2] = x.z; v.l[3] = x.w;
    v.l[4] = ((const uint32_t*)(a + plane))[idx];
  } else {
    v.l[5] = x.x; v.l[6] = x.y; v.l[7] = x.z; v.l[8] = x.w;
  }
}

// ---------------------------------------------------------------- tile

struct TileCtx {
  uint32_t logN, logC, c, i;
  uint32_t in_base, out_base;
  uint32_t plane;  // N * C elements per LDS limb plane
  bool active;
};

// LDS slot of tile element (column t.c, position pos). With SWZ, bits 1-3 of
// the index are flipped by bits 4-6 (a bijection on every aligned group of
// 128 elements, the identity on tiles under 16): the first round of a tile
// with few columns writes elements 16 apart from consecutive threads, which
// land in the same LDS banks unswizzled (a 16-byte slot per element and
// plane, 256 bytes per bank row).
template <bool SWZ>
__device__ __forceinline__ uint32_t lds_elem(const TileCtx& t, uint32_t pos) {
  const uint32_t i = t.c + (pos << t.logC);
  if constexpr (SWZ) return i ^ ((i >> 3) & 0xEu);
  return i;
}

// LDS exchange helpers (PH as in lds_put). A round of radix R = 2^LR at
// stride 2^logp writes its register s*R + bitrev(m) to tile position
// base(s) + m*2^logp; the next round reads its register s*R + j from
// position iv(s) + j*N/R.
template <int E, int R, int PH, bool SWZ>
__device__ __forceinline__ void lds_write_round(const fe (&u)[E], const TileCtx& t, uint32_t logp, uint32_t* lds) {
  constexpr int S = E / R;
  constexpr int LR = Log2<R>::value;
  constexpr int LE = Log2<E>::value;
  const uint32_t NE = 1u << (t.logN - LE);
  const uint32_t pmask = (1u << logp) - 1u;
  static_for<S>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const uint32_t iv = t.i + s * NE;
    const uint32_t k = iv & pmask;
    const uint32_t base = ((iv - k) << LR) + k;
    static_for<R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const uint32_t pos = base + ((uint32_t)m << logp);
      lds_put<PH>(lds, t.plane, lds_elem<SWZ>(t, pos), u[s * R + bitrev_c(m, LR)]);
    });
  });
}
template <int E, int R, int PH, bool SWZ>
__device__ __forceinline__ void lds_read_round(fe (&u)[E], const TileCtx& t, const uint32_t* lds) {
  constexpr int S = E / R;
  constexpr int LR = Log2<R>::value;
  constexpr int LE = Log2<E>::value;
  const uint32_t NE = 1u << (t.logN - LE);
  const uint32_t NR = 1u << (t.logN - LR);
  static_for<S>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const uint32_t iv = t.i + s * NE;
    static_for<R>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      lds_get<PH>(lds, t.plane, lds_elem<SWZ>(t, iv + j * NR), u[s * R + j]);
    });
  });
}

// One Stockham round of radix R on the thread's E registers. A round after
// the first takes its inputs from LDS (exchange X) and multiplies them by the
// round twiddles; the last round writes HBM, the others write their outputs
// (limbs 0-4 of them, with a two-phase exchange) to LDS.
template <int E, int R, class X>
__device__ __forceinline__ void ntt_round(fe (&u)[E], const PassDesc& d, const TileCtx& t, uint32_t logp,
                                          bool first, bool last, uint32_t* lds) {
  constexpr int S = E / R;
  constexpr int LR = Log2<R>::value;
  constexpr int LE = Log2<E>::value;
  const uint32_t NE = 1u << (t.logN - LE);
  const uint32_t NR = 1u << (t.logN - LR);
  const uint32_t pmask = (1u << logp) - 1u;
  const uint4* roots = (const uint4*)d.roots;
  const bool in_planar = d.flags & PASS_IN_PLANAR;

  if (first) {
    if (t.active) {
      static_for<S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const uint32_t iv = t.i + s * NE;
        static_for<R>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          const uint32_t off = t.in_base + (iv + j * NR) * d.in_ts;
          fe v = in_planar ? load_planar(d.in, d.in_plane, off) : load_packed(d.in, off);
          if (d.pre) v = mont_mul<X::chain>(v, load_packed(d.pre, off & d.pre_mask));
          else if (in_planar) reduce_value(v);
          u[s * R + j] = v;
        });
      });
    }
  } else {
    // round twiddles omega_{pR}^{jk}
    const uint32_t sh = d.root_log - (logp + LR);
    auto twiddle = [&](int s, int j) {
      return load_limbs48(roots, (uint64_t)(j * ((t.i + s * NE) & pmask)) << sh);
    };
    if (t.active) lds_read_round<E, R, X::two_phase ? 0 : 2, X::swizzle>(u, t, lds);
    __syncthreads();  // every LDS read precedes the next overwrite
    if constexpr (X::two_phase) {
      // limbs 5-8: the previous round (radix E, since only a tile's last
      // round is shorter) writes them where it wrote limbs 0-4, this round
      // reads them into the registers that now hold its limbs 0-4
      if (t.active) lds_write_round<E, E, 1, X::swizzle>(u, t, logp - LE, lds);
      __syncthreads();
      if (t.active) lds_read_round<E, R, 1, X::swizzle>(u, t, lds);
      __syncthreads();  // every LDS read of this round precedes any overwrite
    }
    if (t.active) {
      static_for<S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        reduce_value(u[s * R]);
        static_for<R, 1>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          u[s * R + j] = mont_mul<X::chain>(u[s * R + j], twiddle(s, j));
        });
      });
    }
  }
  if (t.active) {
    static_for<S>([&](auto sc) { dft_regs<R, X::chain>(&u[decltype(sc)::value * R], roots, d.root_log); });
    static_for<S>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const uint32_t iv = t.i + s * NE;
      const uint32_t k = iv & pmask;
      const uint32_t base = ((iv - k) << LR) + k;
      static_for<R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const uint32_t pos = base + ((uint32_t)m << logp);
        fe v = u[s * R + bitrev_c(m, LR)];
        if (last) {
          const uint32_t off = t.out_base + pos * d.out_ts;
          if (d.flags & PASS_OUT_PLANAR) {
            // DFT outputs meet the multiplier's operand contract; the next
            // pass multiplies them by its pre-twiddle on load
            store_planar(d.out, d.out_plane, off, v);
          } else {
            if (d.post) v = mont_mul<X::chain>(v, load_packed(d.post, off));
            if (d.flags & PASS_USCALE) {
              fe sc;
#pragma unroll
              for (int q = 0; q < NLIMBS; ++q) sc.l[q] = d.scale[q];
              v = mont_mul_u<X::chain>(v, sc);
            }
            uint32_t w[8];
            if (d.flags & PASS_CANON) {
              to_canonical_words(w, v);
            } else {
              reduce_value(v);
              pack_words(w, v);
            }
            store_words(d.out, off, w);
          }
        } else {
          lds_put<X::two_phase ? 0 : 2>(lds, t.plane, lds_elem<X::swizzle>(t, pos), v);
        }
      });
    });
  }
  if (!last) __syncthreads();  // this round's LDS writes precede the next round's reads
}

// Elements per thread: radix-8 rounds keep 8 x 9 data VGPRs per thread
// (LE_MAX = 3); radix-4 rounds keep 4 x 9 (LE_MAX = 2).
template <int LOGN, int LE_MAX = 3> struct TileShape {
  static constexpr int LE = LOGN < LE_MAX ? LOGN : LE_MAX;  // log2 elements per thread
  static constexpr int E = 1 << LE;
  static constexpr int FULL = LOGN / LE;                 // rounds of radix E
  static constexpr int REM = LOGN - FULL * LE;           // log2 radix of the last round (0: none)
  static constexpr int ROUNDS = FULL + (REM ? 1 : 0);
};

// the tile shape of exchange configuration X
template <int LOGN, class X>
using TileOf = TileShape<LOGN, X::le_max>;

template <int LOGN, int RIDX, class X, int E = TileOf<LOGN, X>::E>
__device__ __forceinline__ void ntt_rounds(fe (&u)[E], const PassDesc& d, const TileCtx& t, uint32_t* lds) {
  using Sh = TileOf<LOGN, X>;
  if constexpr (RIDX < Sh::ROUNDS) {
    constexpr int LR = RIDX < Sh::FULL ? Sh::LE : Sh::REM;
    constexpr uint32_t LOGP = (uint32_t)(RIDX * Sh::LE);
    ntt_round<Sh::E, (1 << LR), X>(u, d, t, LOGP, RIDX == 0, RIDX == Sh::ROUNDS - 1, lds);
    ntt_rounds<LOGN, RIDX + 1, X>(u, d, t, lds);
  }
}

// tid: the thread's index within the tile; valid: the tile exists (the
// last group of a batch descriptor may run past its tiles). Inactive threads
// still take every barrier.
template <int LOGN, class X>
__device__ __forceinline__ void ntt_tile(const PassDesc& d, uint32_t tile, uint32_t* lds, uint32_t tid, bool valid) {
  using Sh = TileOf<LOGN, X>;
  TileCtx t;
  t.logN = LOGN;
  t.logC = d.log_c;
  const uint32_t nthreads = 1u << (LOGN + t.logC - Sh::LE);
  t.active = valid && tid < nthreads;
  t.c = tid & ((1u << t.logC) - 1u);
  t.i = tid >> t.logC;
  t.plane = 1u << (LOGN + t.logC);
  const uint32_t tlo = tile & ((1u << d.log_n_lo) - 1u), thi = tile >> d.log_n_lo;
  t.in_base = tlo * d.in_lo + thi * d.in_hi + t.c * d.in_cs;
  t.out_base = tlo * d.out_lo + thi * d.out_hi + t.c * d.out_cs;
  fe u[Sh::E];
  ntt_rounds<LOGN, 0, X>(u, d, t, lds);
}

// Map workgroups so that consecutive tiles land on the same XCD: workgroups
// are dispatched round-robin over the 8 XCDs, and neighbouring tiles share
// HBM lines (adjacent columns) that should meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  return (n & 7u) ? b : (b & 7u) * (n >> 3) + (b >> 3);
}

constexpr int MAX_TILE_THREADS = 1 << (TILE_LOG_MAX - 3);

// Single-pass tiles in three families chosen at launch (quad, below, is the
// third):
//   narrow: tiles up to 2^11 elements (<= 256 threads) in passes with more
//     than two waves per SIMD: two-phase exchange, at most 128 VGPRs, four
//     waves per SIMD (four 40 KiB tiles fill a CU's 160 KiB), single-chain
//     columns;
//   wide: 2^12-element tiles (512 threads; one per CU whatever its LDS, a
//     second would need four waves per SIMD, where these tiles spill), and
//     the tiles of passes too small to give a SIMD more than two waves (any
//     transform below 2^21 elements): the one-phase exchange (two barriers
//     per round fewer) and the two-wave register budget (256 VGPRs; they
//     need 132-137, no spills).
// Both take the single-chain multiplier (one asm chain of multiply-adds per
// column): with two waves per SIMD it still beat the compiler's split
// columns (k=20 step 0.307 vs 0.320 ms), since two waves of eight
// independent products per thread already hide the chain's latency. The
// wide kernel swizzles its LDS slots: its tiles with two columns (k=20 pass
// 2) lost 16 times more cycles to bank conflicts without it; the narrow
// kernel's eight-column tiles conflict little and the swizzle only cost
// instructions there.
constexpr uint32_t NARROW_TILE_LOG = 11;
using PassXchg = Xchg<false, true, 3, true>;
using NarrowXchg = Xchg<true, true>;

template <int LOGN>
__global__ void __launch_bounds__(MAX_TILE_THREADS, 2) ntt_pass_kernel(PassDesc d) {
  extern __shared__ uint32_t lds[];
  ntt_tile<LOGN, PassXchg>(d, xcd_tile(blockIdx.x, d.num_tiles), lds, threadIdx.x, true);
}
template <int LOGN>
__global__ void __launch_bounds__(1 << (NARROW_TILE_LOG - 3), 4) ntt_pass_kernel_narrow(PassDesc d) {
  extern __shared__ uint32_t lds[];
  ntt_tile<LOGN, NarrowXchg>(d, xcd_tile(blockIdx.x, d.num_tiles), lds, threadIdx.x, true);
}
// Quad: the wide configuration with four elements per thread (radix-4
// rounds, a radix-2 round for an odd digit count), for passes that eight
// elements per thread would give fewer than two waves per SIMD (transforms
// below 2^20 elements): twice the waves for more LDS rounds (a 2^19
// transform 0.060 instead of 0.078 ms; at 2^20, with two waves per SIMD
// either way, the extra rounds cost more than the waves gain). 70-78 VGPRs;
// a 2^12 tile is one 1024-thread workgroup.
using QuadXchg = Xchg<false, true, 2, true>;
template <int LOGN>
__global__ void __launch_bounds__(1024, 4) ntt_pass_kernel_quad(PassDesc d) {
  extern __shared__ uint32_t lds[];
  ntt_tile<LOGN, QuadXchg>(d, xcd_tile(blockIdx.x, d.num_tiles), lds, threadIdx.x, true);
}

// Persistent pass kernel for batches: a grid of resident workgroups drains a
// queue of items (any job, any tile length up to 2^BATCH_TILE_LOG elements)
// through an atomic counter. An item is a group of consecutive tiles of one
// descriptor that together fill the workgroup's BATCH_THREADS threads (a
// 2^10-element tile takes 128 threads, so an item holds two of them); each
// tile of the group has its own slice of the LDS planes. Items are ordered by
// descriptor; tile_start is the exclusive prefix sum of the descriptors' item
// counts (batch_items). Every workgroup leaves the loop once the counter
// passes `total`.
//
// LDS layout: the tiles' limb planes (two-phase exchange) from byte 0
// (16-byte aligned, as in the single-pass kernel: the planes are accessed with
// 16-byte DS operations). The claimed item is broadcast through word 0 of
// the planes: when thread 0 writes it, every thread has passed the barrier
// that follows the previous tile's last LDS read, and every thread reads it
// before the barrier that precedes the next tile's first LDS write. So the
// workgroup needs exactly its tile planes, 40 KiB, and four fill a CU.
//
// The persistent kernel inlines every tile length. With the two-phase
// exchange and single-chain columns it fits the four-wave budget of 128
// VGPRs (108 bytes spilled), so four workgroups share a CU (four waves per
// SIMD); at three it needed 157 VGPRs; with the one-phase exchange and split
// columns it needed 234 and ran two.
using BatchXchg = Xchg<true, true, 3, true>;
constexpr uint32_t BATCH_TILE_LOG = 11;  // 2048-element tiles: 40 KiB of LDS
constexpr int BATCH_THREADS = 1 << (BATCH_TILE_LOG - 3);
constexpr uint32_t BATCH_ITEM_WORD = 0;  // LDS word of the item slot (inside the tile planes)

constexpr uint32_t BATCH_THREADS_LOG = BATCH_TILE_LOG - 3;
// resident workgroups per CU (one wave per SIMD each)
constexpr int PERSIST_WGS = 4;

// Item `group` of descriptor d: tiles group * G .. group * G + G - 1, G =
// BATCH_THREADS / (threads per tile).
template <int L>
__device__ __forceinline__ void tile_dispatch(const PassDesc& d, uint32_t group, uint32_t* lds) {
  if constexpr (L <= (int)BATCH_TILE_LOG) {
    if (d.log_n_tile == (uint32_t)L) {
      const uint32_t lgt = L + d.log_c - TileOf<L, BatchXchg>::LE;  // log2 threads per tile
      const uint32_t sub = threadIdx.x >> lgt;
      const uint32_t tile = (group << (BATCH_THREADS_LOG - lgt)) + sub;
      ntt_tile<L, BatchXchg>(d, tile, lds + sub * (BatchXchg::words << (L + d.log_c)),
                             threadIdx.x & ((1u << lgt) - 1u), tile < d.num_tiles);
      return;
    }
    tile_dispatch<L + 1>(d, group, lds);
  }
}

__global__ void __launch_bounds__(BATCH_THREADS, PERSIST_WGS) ntt_persistent_kernel(const PassDesc* __restrict__ descs,
                                                                       const uint32_t* __restrict__ tile_start,
                                                                       uint32_t ndesc, uint32_t total,
                                                                       uint32_t* counter) {
  extern __shared__ __align__(16) uint32_t lds[];
  for (;;) {
    // every LDS read of the previous tile, and of the slot's previous value,
    // precedes a barrier that thread 0 has passed
    if (threadIdx.x == 0) lds[BATCH_ITEM_WORD] = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t item = lds[BATCH_ITEM_WORD];
    __syncthreads();  // every thread has read the slot before the tile reuses LDS
    if (item >= total) break;
    uint32_t lo = 0, hi = ndesc;  // largest idx with tile_start[idx] <= item
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (tile_start[mid] <= item) lo = mid; else hi = mid;
    }
    const PassDesc d = descs[lo];
    tile_dispatch<1>(d, item - tile_start[lo], lds);
  }
}

__global__ void reset_counters_kernel(uint32_t* c, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0;
}

// ---------------------------------------------------------------- tables

struct FeArg { uint32_t w[8]; };

// x * 2^256 (packed words) -> x * 2^261, normalized and canonical
__device__ __forceinline__ fe to_domain261(const uint32_t w[8]) {
  fe v = mont_mul(unpack_words(w), fe_from_const(TO261));
  uint32_t c[8];
  to_canonical_words(c, v);
  return unpack_words(c);
}
__device__ __forceinline__ fe canon_fe(const fe& v) {
  uint32_t c[8];
  to_canonical_words(c, v);
  return unpack_words(c);
}

// out[e] = base^(e * step) in the 2^261 domain, e < count (square and multiply).
// packed48 != 0: 48-byte limb entries; otherwise packed 32-byte words.
__global__ void pow_table_kernel(void* out, FeArg base, uint64_t step, uint32_t count, int limbs48) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  fe sq = to_domain261(base.w);
  fe acc = fe_from_const(ONE261);
  uint64_t x = (uint64_t)e * step;
  while (x) {
    if (x & 1) acc = canon_fe(mont_mul(acc, sq));
    x >>= 1;
    if (x) sq = canon_fe(mont_mul(sq, sq));
  }
  if (limbs48) {
    store_limbs48(out, e, acc);
  } else {
    uint32_t w[8];
    to_canonical_words(w, acc);
    store_words(out, e, w);
  }
}

// out[e] = lo[e & (2^h-1)] * hi[e >> h] (* scale), e < count; lo, hi in 48-byte
// limb form (2^261 domain); out packed words in the 2^261 domain.
__global__ void pow_table_2level_kernel(void* out, const void* lo, const void* hi, uint32_t h, uint64_t count,
                                        FeArg scale, int use_scale) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  fe v = mont_mul(load_limbs48(lo, e & ((1ull << h) - 1)), load_limbs48(hi, e >> h));
  if (use_scale) v = mont_mul(canon_fe(v), to_domain261(scale.w));
  uint32_t w[8];
  to_canonical_words(w, v);
  store_words(out, e, w);
}

// Inter-pass twiddles of a sub-problem of size m = 2^log_m split as N x L:
// out[t * L + i] = w_m^((t * i) mod m) (* scale), via w_m^e = lo[e & (2^h-1)] *
// hi[e >> h]. The scale folds a uniform factor (the inverse's divisor) into
// the twiddles the last pass applies to every element on load.
__global__ void twiddle_table_kernel(void* out, const void* lo, const void* hi, uint32_t h, uint32_t log_m,
                                     uint32_t log_L, FeArg scale, int use_scale) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (1ull << log_m)) return;
  const uint64_t t = idx >> log_L, i = idx & ((1ull << log_L) - 1);
  const uint64_t e = (t * i) & ((1ull << log_m) - 1);
  fe v = mont_mul(load_limbs48(lo, e & ((1ull << h) - 1)), load_limbs48(hi, e >> h));
  if (use_scale) v = mont_mul(canon_fe(v), to_domain261(scale.w));
  uint32_t w[8];
  to_canonical_words(w, v);
  store_words(out, idx, w);
}

// Seeded synthetic elements (Montgomery form x * 2^256, x uniform in [0, r)):
// element i tries attempts a = 0, 1, ...; attempt a takes four SplitMix64
// outputs from the state seed * 0xD1B54A32D192ED03 + (i * 64 + a) * 4, masks
// the top limb to 62 bits and accepts the value when it is below r.
__device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ bool words_below_r(const uint32_t w[8]) {
  for (int j = 7; j >= 0; --j) {
    if (w[j] < p32(j)) return true;
    if (w[j] > p32(j)) return false;
  }
  return false;
}
__global__ void fill_random_kernel(void* out, uint64_t seed, uint64_t start, uint64_t count) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint64_t i = start + k;
  const uint64_t gamma = 0x9E3779B97F4A7C15ull;
  uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t a = 0; a < 64; ++a) {
    const uint64_t st = seed * 0xD1B54A32D192ED03ull + (i * 64 + a) * 4;
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint64_t l = splitmix_mix(st + (uint64_t)(j + 1) * gamma);
      if (j == 3) l &= (1ull << 62) - 1;
      v[2 * j] = (uint32_t)l;
      v[2 * j + 1] = (uint32_t)(l >> 32);
    }
    if (words_below_r(v)) {
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = v[j];
      break;
    }
  }
  const fe v = mont_mul(unpack_words(x), fe_from_const(ENCODE256));
  uint32_t w[8];
  to_canonical_words(w, v);
  store_words(out, k, w);
}

// flag[0] |= 1 when some element is >= r (non-canonical input)
__global__ void check_canonical_kernel(const void* in, uint64_t count, uint32_t* flag) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint4* p = (const uint4*)in + 2 * k;
  const uint4 lo = p[0], hi = p[1];
  const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  if (!words_below_r(w)) atomicOr(flag, 1u);
}

// ---------------------------------------------------------------- launchers

uint32_t tile_threads(uint32_t log_n_tile, uint32_t log_c, uint32_t le_max) {
  const uint32_t le = log_n_tile < le_max ? log_n_tile : le_max;
  return 1u << (log_n_tile + log_c - le);
}
// a pass runs the narrow family when it has more waves than this (two per
// SIMD of the device; set by configure_kernels)
static std::atomic<uint64_t> g_narrow_min_waves{2048};
// queue items of a batch descriptor: groups of tiles that fill BATCH_THREADS
uint32_t batch_items(const PassDesc& d) {
  const uint32_t per_item = (uint32_t)BATCH_THREADS / tile_threads(d.log_n_tile, d.log_c, 3);
  return (d.num_tiles + per_item - 1) / per_item;
}
uint32_t persistent_wgs_per_cu() { return (uint32_t)PERSIST_WGS; }
// planes of the largest batch tile (the item slot lies inside them)
static uint32_t persistent_lds_bytes() { return (4u * BatchXchg::words) << BATCH_TILE_LOG; }

// Which pass-kernel family runs a descriptor (see ntt_pass_kernel_narrow).
template <int LOGN>
static bool use_narrow(const PassDesc& d) {
  if constexpr (LOGN <= (int)NARROW_TILE_LOG) {
    const uint32_t tl = LOGN + d.log_c;
    const uint32_t threads = tile_threads(LOGN, d.log_c, NarrowXchg::le_max);
    const uint64_t waves = ((uint64_t)d.num_tiles * threads + 63) / 64;
    return tl <= NARROW_TILE_LOG && waves > g_narrow_min_waves.load(std::memory_order_relaxed);
  }
  return false;
}

// Passes left to the wide configuration that eight elements per thread would
// give fewer than two waves per SIMD run the quad configuration.
template <int LOGN>
static bool use_quad(const PassDesc& d) {
  if (use_narrow<LOGN>(d)) return false;
  const uint32_t threads = tile_threads(LOGN, d.log_c, PassXchg::le_max);
  const uint64_t waves = ((uint64_t)d.num_tiles * threads + 63) / 64;
  return waves < g_narrow_min_waves.load(std::memory_order_relaxed);
}

// <<<>>> when no event is asked for; otherwise hipExtLaunchKernel, which ties
// the events to the dispatch itself (no separate event packet on the stream).
template <class K>
static hipError_t launch_kernel(K kernel, uint32_t grid, uint32_t threads, uint32_t lds, hipStream_t s,
                                const PassDesc& d, hipEvent_t start, hipEvent_t stop) {
  if (!start && !stop) {
    kernel<<<grid, threads, lds, s>>>(d);
    return hipGetLastError();
  }
  PassDesc arg = d;
  void* args[] = {&arg};
  return hipExtLaunchKernel((const void*)kernel, dim3(grid), dim3(threads), args, lds, s, start, stop, 0);
}

template <int LOGN>
static hipError_t launch_one(const PassDesc& d, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const uint32_t tl = LOGN + d.log_c;
  if constexpr (LOGN <= (int)NARROW_TILE_LOG) {
    if (use_narrow<LOGN>(d)) {
      return launch_kernel(ntt_pass_kernel_narrow<LOGN>, d.num_tiles, tile_threads(LOGN, d.log_c, NarrowXchg::le_max),
                           (4u * NarrowXchg::words) << tl, s, d, start, stop);
    }
  }
  if (use_quad<LOGN>(d)) {
    return launch_kernel(ntt_pass_kernel_quad<LOGN>, d.num_tiles, tile_threads(LOGN, d.log_c, QuadXchg::le_max),
                         (4u * QuadXchg::words) << tl, s, d, start, stop);
  }
  return launch_kernel(ntt_pass_kernel<LOGN>, d.num_tiles, tile_threads(LOGN, d.log_c, PassXchg::le_max),
                       (4u * PassXchg::words) << tl, s, d, start, stop);
}
#define ZK_TILE_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

hipError_t launch_pass(const PassDesc& d, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  switch (d.log_n_tile) {
#define X(L) case L: return launch_one<L>(d, s, start, stop);
    ZK_TILE_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

bool pass_kernel_quad(const PassDesc& d) {
  switch (d.log_n_tile) {
#define X(L) case L: return use_quad<L>(d);
    ZK_TILE_CASES(X)
#undef X
    default: return false;
  }
}

bool pass_kernel_narrow(const PassDesc& d) {
  switch (d.log_n_tile) {
#define X(L) case L: return use_narrow<L>(d);
    ZK_TILE_CASES(X)
#undef X
    default: return false;
  }
}

hipError_t launch_persistent(const PassDesc* d_descs, const uint32_t* d_tile_start, uint32_t ndesc, uint32_t total,
                             uint32_t* d_counter, uint32_t grid, hipStream_t s) {
  ntt_persistent_kernel<<<grid, BATCH_THREADS, persistent_lds_bytes(), s>>>(d_descs, d_tile_start, ndesc,
                                                                                     total, d_counter);
  return hipGetLastError();
}
hipError_t launch_reset_counters(uint32_t* d_counter, uint32_t n, hipStream_t s) {
  reset_counters_kernel<<<(n + 63) / 64, 64, 0, s>>>(d_counter, n);
  return hipGetLastError();
}
uint32_t batch_tile_log() { return BATCH_TILE_LOG; }

static FeArg to_arg(const uint32_t v[8]) {
  FeArg a;
  for (int i = 0; i < 8; ++i) a.w[i] = v[i];
  return a;
}

hipError_t launch_pow_table(void* out, const uint32_t base[8], uint64_t step, uint32_t count, bool limbs48,
                            hipStream_t s) {
  pow_table_kernel<<<(count + 255) / 256, 256, 0, s>>>(out, to_arg(base), step, count, limbs48 ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_pow_table_2level(void* out, const void* lo, const void* hi, uint32_t h, uint64_t count,
                                   const uint32_t* scale, hipStream_t s) {
  uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  pow_table_2level_kernel<<<(uint32_t)((count + 255) / 256), 256, 0, s>>>(
      out, lo, hi, h, count, to_arg(scale ? scale : zero), scale ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_twiddle_table(void* out, const void* lo, const void* hi, uint32_t h, uint32_t log_m,
                                uint32_t log_L, const uint32_t* scale, hipStream_t s) {
  const uint64_t m = 1ull << log_m;
  uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  twiddle_table_kernel<<<(uint32_t)((m + 255) / 256), 256, 0, s>>>(out, lo, hi, h, log_m, log_L,
                                                                  to_arg(scale ? scale : zero), scale ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_fill_random(void* out, uint64_t seed, uint64_t start, uint64_t count, hipStream_t s) {
  fill_random_kernel<<<(uint32_t)((count + 255) / 256), 256, 0, s>>>(out, seed, start, count);
  return hipGetLastError();
}
hipError_t launch_check_canonical(const void* in, uint64_t count, uint32_t* flag, hipStream_t s) {
  check_canonical_kernel<<<(uint32_t)((count + 255) / 256), 256, 0, s>>>(in, count, flag);
  return hipGetLastError();
}

// Throughput probe of the 32x32->64 multiply-add (v_mad_u64_u32), the
// instruction the Montgomery products are made of: 8 independent chains per
// thread, each step feeding the accumulator's high word back as a factor so
// that nothing can be folded; registers only, one store per thread.
constexpr int MAD_PROBE_CHAINS = 8;
constexpr int MAD_PROBE_ITERS = 2048;
__global__ void __launch_bounds__(256) mad_probe_kernel(uint64_t* out, uint32_t seed) {
  uint64_t acc[MAD_PROBE_CHAINS];
  uint32_t mulv[MAD_PROBE_CHAINS];
#pragma unroll
  for (int c = 0; c < MAD_PROBE_CHAINS; ++c) {
    acc[c] = seed + threadIdx.x * 977u + c;
    mulv[c] = 0x9E3779B9u * (c + 1) + seed;
  }
  for (int it = 0; it < MAD_PROBE_ITERS; ++it) {
#pragma unroll
    for (int c = 0; c < MAD_PROBE_CHAINS; ++c) acc[c] = (uint64_t)(uint32_t)(acc[c] >> 32) * mulv[c] + acc[c];
  }
  uint64_t x = 0;
#pragma unroll
  for (int c = 0; c < MAD_PROBE_CHAINS; ++c) x ^= acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}
uint64_t mad_probe_ops_per_thread() { return (uint64_t)MAD_PROBE_CHAINS * MAD_PROBE_ITERS; }
hipError_t launch_mad_probe(uint64_t* out, uint32_t blocks, uint32_t seed, hipStream_t s) {
  mad_probe_kernel<<<blocks, 256, 0, s>>>(out, seed);
  return hipGetLastError();
}

hipError_t configure_kernels() {
  const int max_lds = (int)(4u * PassXchg::words) << TILE_LOG_MAX;
  int dev = 0, cus = 256;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) g_narrow_min_waves.store(8ull * (uint64_t)cus, std::memory_order_relaxed);  // 2 x 4 SIMDs
#define X(L)                                                                                              \
  if (e == hipSuccess)                                                                                    \
    e = hipFuncSetAttribute((const void*)ntt_pass_kernel<L>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            max_lds);                                                                     \
  if (e == hipSuccess)                                                                                    \
    e = hipFuncSetAttribute((const void*)ntt_pass_kernel_quad<L>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)(4u * QuadXchg::words) << TILE_LOG_MAX);                                 \
  if (e == hipSuccess && L <= (int)NARROW_TILE_LOG)                                                       \
    e = hipFuncSetAttribute((const void*)ntt_pass_kernel_narrow<(L <= 11 ? L : 11)>,                      \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                                   \
                            (int)(4u * NarrowXchg::words) << NARROW_TILE_LOG);                            \
  ZK_TILE_CASES(X)
#undef X
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)ntt_persistent_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)persistent_lds_bytes());
  return e;
}

}  // namespace zkgpu
