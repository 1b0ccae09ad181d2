// HIP kernels of the BN256 Fr NTT for gfx950 (MI355X).
//
// ntt_tile: one workgroup owns C columns x N positions of one pass (see
// ntt_desc.hpp). Each thread keeps 8 elements in 29-bit-limb form
// (8 x 9 VGPRs) and the tile runs Stockham (self-sorting) rounds of radix
// 8 (the last round takes the remaining radix 2/4): round 1 reads HBM,
// middle rounds exchange through LDS (9 dword planes, conflict-free for
// consecutive positions), the last round writes HBM. Every pass therefore
// reads and writes each element once. Inside a round a thread multiplies its
// inputs by the round twiddles omega_{pR}^{jk}, runs a radix-2
// decimation-in-frequency network over its R registers and scatters the
// (bit-reversal relabelled) outputs.
//
// Pass-level fusions: the coset pre-scale (a[i] *= c^i, reference
// cuda/fft.cu:223-232) on the first load; the inter-pass twiddle, the
// inverse's divisor (reference cuda/fft.cu:153-155) and the coset-inverse
// post-scale on the last store, together with the final canonicalisation.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

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
// unpacked table entry: 9 limbs in 12 u32 (48 bytes)
__device__ __forceinline__ fe load_limbs48(const void* base, uint64_t idx) {
  const uint4* p = (const uint4*)base + 3 * idx;
  const uint4 a = p[0], b = p[1], c = p[2];
  fe r;
  r.l[0] = a.x; r.l[1] = a.y; r.l[2] = a.z; r.l[3] = a.w;
  r.l[4] = b.x; r.l[5] = b.y; r.l[6] = b.z; r.l[7] = b.w;
  r.l[8] = c.x;
  return r;
}
__device__ __forceinline__ void store_limbs48(void* base, uint64_t idx, const fe& v) {
  uint4* p = (uint4*)base + 3 * idx;
  p[0] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  p[1] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
  p[2] = make_uint4(v.l[8], 0u, 0u, 0u);
}
// LDS: nine planes of `plane` dwords
__device__ __forceinline__ fe lds_load(const uint32_t* lds, uint32_t idx, uint32_t plane) {
  fe r;
#pragma unroll
  for (int q = 0; q < NLIMBS; ++q) r.l[q] = lds[q * plane + idx];
  return r;
}
__device__ __forceinline__ void lds_store(uint32_t* lds, uint32_t idx, uint32_t plane, const fe& v) {
#pragma unroll
  for (int q = 0; q < NLIMBS; ++q) lds[q * plane + idx] = v.l[q];
}

template <int R>
struct Log2 { static constexpr int value = R <= 1 ? 0 : 1 + Log2<R / 2>::value; };

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, N). Keeps
// register-array indices constant without relying on the unroller.
template <int N, int B = 0, class F>
__device__ __forceinline__ void static_for(F&& f) {
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

// Value bounds (units of r) of the radix-2 DIF network: inputs < 3r; stage s
// outputs are sums (<= 2B), multiplier outputs (< 1.5r) or, for the unit
// twiddle, differences x + C_s r - y (<= 2B + 2). B: 3, 8, 18, 38, 78.
// C_s = B_s + 2 makes C_s r exceed every y of stage s, as sub<C> requires.
template <int S> struct StageC { static constexpr uint32_t value = 2 * StageC<S - 1>::value; };
template <> struct StageC<0> { static constexpr uint32_t value = 5; };

template <int S, int R, int H>
__device__ __forceinline__ void dif_stage(fe* u, const uint4* roots, uint32_t root_log) {
  constexpr int LR = Log2<R>::value;
  constexpr uint32_t C = StageC<S>::value;
  static_for<H>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    fe w;
    if constexpr (j != 0) w = load_limbs48(roots, (uint64_t)(j * (R / (2 * H))) << (root_log - LR));
    static_for<R / (2 * H)>([&](auto bc) {
      constexpr int b = decltype(bc)::value * 2 * H;
      const fe x = u[b + j], y = u[b + j + H];
      fe s = add(x, y);
      normalize(s);
      fe d = sub<C>(x, y);
      if constexpr (j == 0) normalize(d);
      else d = mont_mul(d, w);
      u[b + j] = s;
      u[b + j + H] = d;
    });
  });
}

// Length-R DFT in registers: natural order in, bit-reversed order out.
// omega_R^e = roots[e << (root_log - log R)]. Inputs normalized, V < 3r;
// outputs normalized, V <= 78r.
template <int R>
__device__ __forceinline__ void dft_regs(fe* u, const uint4* roots, uint32_t root_log) {
  if constexpr (R >= 16) dif_stage<0, R, 8>(u, roots, root_log);
  if constexpr (R >= 8) dif_stage<(R >= 16 ? 1 : 0), R, 4>(u, roots, root_log);
  if constexpr (R >= 4) dif_stage<(R >= 16 ? 2 : R >= 8 ? 1 : 0), R, 2>(u, roots, root_log);
  dif_stage<(R >= 16 ? 3 : R >= 8 ? 2 : R >= 4 ? 1 : 0), R, 1>(u, roots, root_log);
}

// ---------------------------------------------------------------- tile

struct TileCtx {
  uint32_t logN, logC, c, i;
  uint32_t in_base, out_base, w_base;
  uint32_t plane;  // N * C dwords per LDS limb plane
  bool active;
};

template <int E, int R>
__device__ __forceinline__ void ntt_round(fe (&u)[E], const PassDesc& d, const TileCtx& t, uint32_t logp,
                                          bool first, bool last, uint32_t* lds) {
  constexpr int S = E / R;
  constexpr int LR = Log2<R>::value;
  constexpr int LE = Log2<E>::value;
  const uint32_t NE = 1u << (t.logN - LE);
  const uint32_t NR = 1u << (t.logN - LR);
  const uint32_t pmask = (1u << logp) - 1u;
  const uint4* roots = (const uint4*)d.roots;

  if (t.active) {
    static_for<S>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const uint32_t iv = t.i + s * NE;
      static_for<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const uint32_t pos = iv + j * NR;
        if (first) {
          const uint32_t off = t.in_base + pos * d.in_ts;
          fe v = load_packed(d.in, off);
          if (d.pre) v = mont_mul(v, load_packed(d.pre, off));
          u[s * R + j] = v;
        } else {
          u[s * R + j] = lds_load(lds, t.c + (pos << t.logC), t.plane);
        }
      });
    });
  }
  if (!first) __syncthreads();  // every LDS read of this round precedes any overwrite
  if (t.active) {
    if (logp > 0) {
      const uint32_t sh = d.root_log - (logp + LR);
      static_for<S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const uint32_t k = (t.i + s * NE) & pmask;
        reduce_value(u[s * R]);
        static_for<R, 1>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          u[s * R + j] = mont_mul(u[s * R + j], load_limbs48(roots, (uint64_t)(j * k) << sh));
        });
      });
    }
    static_for<S>([&](auto sc) { dft_regs<R>(&u[decltype(sc)::value * R], roots, d.root_log); });
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
          if (d.wtab) v = mont_mul(v, load_packed(d.wtab, t.w_base + pos * d.w_ts));
          const uint32_t off = t.out_base + pos * d.out_ts;
          if (d.post) v = mont_mul(v, load_packed(d.post, off));
          if (d.flags & PASS_USCALE) {
            fe sc;
#pragma unroll
            for (int q = 0; q < NLIMBS; ++q) sc.l[q] = d.scale[q];
            v = mont_mul(v, sc);
          }
          uint32_t w[8];
          if (d.flags & PASS_CANON) {
            to_canonical_words(w, v);
          } else {
            reduce_value(v);
            pack_words(w, v);
          }
          store_words(d.out, off, w);
        } else {
          lds_store(lds, t.c + (pos << t.logC), t.plane, v);
        }
      });
    });
  }
  if (!last) __syncthreads();  // this round's LDS writes precede the next round's reads
}

// Elements per thread: radix-8 rounds keep 8 x 9 data VGPRs per thread.
template <int LOGN> struct TileShape {
  static constexpr int LE = LOGN < 3 ? LOGN : 3;         // log2 elements per thread
  static constexpr int E = 1 << LE;
  static constexpr int FULL = LOGN / LE;                 // rounds of radix E
  static constexpr int REM = LOGN - FULL * LE;           // log2 radix of the last round (0: none)
  static constexpr int ROUNDS = FULL + (REM ? 1 : 0);
};

template <int LOGN, int RIDX>
__device__ __forceinline__ void ntt_rounds(fe (&u)[TileShape<LOGN>::E], const PassDesc& d, const TileCtx& t,
                                           uint32_t* lds) {
  using Sh = TileShape<LOGN>;
  if constexpr (RIDX < Sh::ROUNDS) {
    constexpr int LR = RIDX < Sh::FULL ? Sh::LE : Sh::REM;
    constexpr uint32_t LOGP = (uint32_t)(RIDX * Sh::LE);
    ntt_round<Sh::E, (1 << LR)>(u, d, t, LOGP, RIDX == 0, RIDX == Sh::ROUNDS - 1, lds);
    ntt_rounds<LOGN, RIDX + 1>(u, d, t, lds);
  }
}

template <int LOGN>
__device__ __forceinline__ void ntt_tile(const PassDesc& d, uint32_t tile, uint32_t* lds) {
  using Sh = TileShape<LOGN>;
  TileCtx t;
  t.logN = LOGN;
  t.logC = d.log_c;
  const uint32_t nthreads = 1u << (LOGN + t.logC - Sh::LE);
  const uint32_t tid = threadIdx.x;
  t.active = tid < nthreads;
  t.c = tid & ((1u << t.logC) - 1u);
  t.i = tid >> t.logC;
  t.plane = 1u << (LOGN + t.logC);
  const uint32_t tlo = tile & ((1u << d.log_n_lo) - 1u), thi = tile >> d.log_n_lo;
  t.in_base = tlo * d.in_lo + thi * d.in_hi + t.c * d.in_cs;
  t.out_base = tlo * d.out_lo + thi * d.out_hi + t.c * d.out_cs;
  t.w_base = tlo * d.w_lo + thi * d.w_hi + t.c * d.w_cs;
  fe u[Sh::E];
  ntt_rounds<LOGN, 0>(u, d, t, lds);
}

// Map workgroups so that consecutive tiles land on the same XCD: workgroups
// are dispatched round-robin over the 8 XCDs, and neighbouring tiles share
// HBM lines (adjacent columns) that should meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  return (n & 7u) ? b : (b & 7u) * (n >> 3) + (b >> 3);
}

constexpr int MAX_TILE_THREADS = 1 << (TILE_LOG_MAX - 3);

template <int LOGN>
__global__ void __launch_bounds__(MAX_TILE_THREADS) ntt_pass_kernel(PassDesc d) {
  extern __shared__ uint32_t lds[];
  ntt_tile<LOGN>(d, xcd_tile(blockIdx.x, d.num_tiles), lds);
}

// Many jobs' passes in one launch, all with tile length 2^LOGN: tile_start is
// the exclusive prefix sum of the descriptors' tile counts; blockDim is the
// largest tile's thread count (smaller tiles idle their extra threads).
template <int LOGN>
__global__ void __launch_bounds__(MAX_TILE_THREADS) ntt_pass_batched_kernel(const PassDesc* __restrict__ descs,
                                                                           const uint32_t* __restrict__ tile_start,
                                                                           uint32_t ndesc, uint32_t total_tiles) {
  extern __shared__ uint32_t lds[];
  const uint32_t b = xcd_tile(blockIdx.x, total_tiles);
  uint32_t lo = 0, hi = ndesc;  // largest idx with tile_start[idx] <= b
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tile_start[mid] <= b) lo = mid; else hi = mid;
  }
  const PassDesc d = descs[lo];
  ntt_tile<LOGN>(d, b - tile_start[lo], lds);
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
// out[t * L + i] = w_m^((t * i) mod m), via w_m^e = lo[e & (2^h-1)] * hi[e >> h].
__global__ void twiddle_table_kernel(void* out, const void* lo, const void* hi, uint32_t h, uint32_t log_m,
                                     uint32_t log_L) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (1ull << log_m)) return;
  const uint64_t t = idx >> log_L, i = idx & ((1ull << log_L) - 1);
  const uint64_t e = (t * i) & ((1ull << log_m) - 1);
  const fe v = mont_mul(load_limbs48(lo, e & ((1ull << h) - 1)), load_limbs48(hi, e >> h));
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

uint32_t tile_threads(uint32_t log_n_tile, uint32_t log_c) {
  const uint32_t le = log_n_tile < 3 ? log_n_tile : 3;
  return 1u << (log_n_tile + log_c - le);
}
uint32_t tile_lds_bytes(uint32_t log_n_tile, uint32_t log_c) {
  return (uint32_t)(4u * NLIMBS) << (log_n_tile + log_c);
}

template <int LOGN>
static void launch_one(const PassDesc& d, hipStream_t s) {
  ntt_pass_kernel<LOGN><<<d.num_tiles, tile_threads(LOGN, d.log_c), tile_lds_bytes(LOGN, d.log_c), s>>>(d);
}
template <int LOGN>
static void launch_batched_one(const PassDesc* descs, const uint32_t* ts, uint32_t ndesc, uint32_t total,
                               uint32_t threads, uint32_t lds, hipStream_t s) {
  ntt_pass_batched_kernel<LOGN><<<total, threads, lds, s>>>(descs, ts, ndesc, total);
}

#define ZK_TILE_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

hipError_t launch_pass(const PassDesc& d, hipStream_t s) {
  switch (d.log_n_tile) {
#define X(L) case L: launch_one<L>(d, s); break;
    ZK_TILE_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// All descriptors of one batched launch share log_n_tile.
hipError_t launch_pass_batched(uint32_t log_n_tile, const PassDesc* d_descs, const uint32_t* d_tile_start,
                               uint32_t ndesc, uint32_t total_tiles, uint32_t block_threads, uint32_t lds_bytes,
                               hipStream_t s) {
  switch (log_n_tile) {
#define X(L) case L: launch_batched_one<L>(d_descs, d_tile_start, ndesc, total_tiles, block_threads, lds_bytes, s); break;
    ZK_TILE_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

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
                                uint32_t log_L, hipStream_t s) {
  const uint64_t m = 1ull << log_m;
  twiddle_table_kernel<<<(uint32_t)((m + 255) / 256), 256, 0, s>>>(out, lo, hi, h, log_m, log_L);
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

hipError_t configure_kernels() {
  const int max_lds = (int)(4u * NLIMBS) << TILE_LOG_MAX;
  hipError_t e = hipSuccess;
#define X(L)                                                                                              \
  if (e == hipSuccess)                                                                                    \
    e = hipFuncSetAttribute((const void*)ntt_pass_kernel<L>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            max_lds);                                                                     \
  if (e == hipSuccess)                                                                                    \
    e = hipFuncSetAttribute((const void*)ntt_pass_batched_kernel<L>,                                      \
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
  ZK_TILE_CASES(X)
#undef X
  return e;
}

}  // namespace zkgpu
