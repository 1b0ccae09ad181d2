// HIP kernels of the G1 multi-scalar multiplication (Pippenger bucket method).
//
// sum_i s_i * P_i with s_i in Fr (halo2 Montgomery words, read as canonical
// integers as the reference does, cuda/curve.cuh:75-99) and P_i affine points
// over Fq. Scalars are cut into W windows of c bits, recoded as signed digits
// in [-2^(c-1), 2^(c-1)], so each window needs 2^(c-1) buckets and a negative
// digit adds -P_i. The reference instead gives each thread a run of points
// and 127 buckets in its own stack (cuda/fft.cu:158-202); here the buckets of
// all points are shared:
//
//   msm_digits    per point: base to the 2^261 domain, scalar to signed digits
//                 (one 16-bit key per window)
//   sort          the (window, point) pairs into bucket order: two radix
//                 passes through LDS (partition scan, partition, then per
//                 slice of a partition bucket counts and scatter)
//   msm_buckets   one thread per bucket segment: sum of its points (mixed
//                 additions in XYZZ coordinates); then a pairwise tree of
//                 the segment sums of buckets that have several
//   msm_chunks    two threads per run of L buckets of a window: running sums
//                 give sum_j (j+1) B_j over the run
//   msm_windows   one workgroup per window: tree sum of its runs, stored as
//                 canonical halo2 words
//
// The windows are combined (Horner, c doublings each) by the host.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fq29.hpp"

namespace zkgpu {

// a base as the bucket sums read it: x and y in the 2^261 domain as
// canonical packed words (below q < 2^254), 64 bytes, so that each random
// read is one aligned 64-byte segment; bit 31 of y's top word marks the
// identity
struct BaseLimbs {
  uint4 x0, x1, y0, y1;
};
// a bucket / partial sum: XYZZ coordinates X, Y, ZZ, ZZZ (nine limbs each) in
// 144-byte entries
struct PointLimbs {
  uint4 a[9];
};

__device__ __forceinline__ fe fe_from3(const uint4& a, const uint4& b, const uint4& c) {
  fe r;
  r.l[0] = a.x; r.l[1] = a.y; r.l[2] = a.z; r.l[3] = a.w;
  r.l[4] = b.x; r.l[5] = b.y; r.l[6] = b.z; r.l[7] = b.w;
  r.l[8] = c.x;
  return r;
}
__device__ __forceinline__ void fe_to3(const fe& v, uint4& a, uint4& b, uint4& c) {
  a = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  b = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
  c = make_uint4(v.l[8], 0u, 0u, 0u);
}
// the base's coordinates as normalized limbs, value below q
__device__ __forceinline__ g1a unpack_base(const BaseLimbs& e) {
  const uint32_t xw[8] = {e.x0.x, e.x0.y, e.x0.z, e.x0.w, e.x1.x, e.x1.y, e.x1.z, e.x1.w};
  const uint32_t yw[8] = {e.y0.x, e.y0.y, e.y0.z, e.y0.w, e.y1.x, e.y1.y, e.y1.z, e.y1.w & 0x7fffffffu};
  g1a p;
  p.x = unpack_words(xw);
  p.y = unpack_words(yw);
  p.inf = (e.y1.w >> 31) != 0;
  return p;
}
__device__ __forceinline__ g1a load_base(const BaseLimbs* b, uint32_t i) { return unpack_base(b[i]); }
__device__ __forceinline__ g1xyzz load_point(const PointLimbs* p, uint32_t i) {
  const PointLimbs e = p[i];
  uint32_t w[36];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    w[4 * k] = e.a[k].x;
    w[4 * k + 1] = e.a[k].y;
    w[4 * k + 2] = e.a[k].z;
    w[4 * k + 3] = e.a[k].w;
  }
  g1xyzz v;
#pragma unroll
  for (int k = 0; k < NLIMBS; ++k) {
    v.x.l[k] = w[k];
    v.y.l[k] = w[NLIMBS + k];
    v.zz.l[k] = w[2 * NLIMBS + k];
    v.zzz.l[k] = w[3 * NLIMBS + k];
  }
  return v;
}
__device__ __forceinline__ void store_point(PointLimbs* p, uint32_t i, const g1xyzz& v) {
  uint32_t w[36];
#pragma unroll
  for (int k = 0; k < NLIMBS; ++k) {
    w[k] = v.x.l[k];
    w[NLIMBS + k] = v.y.l[k];
    w[2 * NLIMBS + k] = v.zz.l[k];
    w[3 * NLIMBS + k] = v.zzz.l[k];
  }
  PointLimbs e;
#pragma unroll
  for (int k = 0; k < 9; ++k) e.a[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
  p[i] = e;
}

// c-bit field of the 256-bit canonical scalar at bit `pos` (c <= 16)
__device__ __forceinline__ uint32_t scalar_bits(const uint32_t w[8], uint32_t pos, uint32_t c) {
  const uint32_t q = pos >> 5, s = pos & 31;
  uint64_t v = q < 8 ? w[q] : 0;
  if (q + 1 < 8) v |= (uint64_t)w[q + 1] << 32;
  return (uint32_t)(v >> s) & ((1u << c) - 1u);
}

// 2^5 in radix 2^29: mont_mul(x * 2^256, 32) = x (halo2 Fr words -> canonical)
__device__ __forceinline__ fe fr_from_mont_factor() {
  fe f;
#pragma unroll
  for (int i = 0; i < NLIMBS; ++i) f.l[i] = 0;
  f.l[0] = 32;
  return f;
}

// Sorting the (window, point) pairs by bucket. A pair's key is the bucket
// |d| - 1 of its digit d (c - 1 bits) with the sign in bit 15, MSM_KEY_NONE
// for a zero digit (or an identity base). The bucket splits into hi = key >>
// lo_bits (the partition, up to MSM_MAX_PARTS per window) and lo (the bucket
// within its partition, lo_bits <= 8). Two radix passes, each staged through
// LDS so that writes leave in runs:
//   msm_digits_kernel       keys, and per-window partition sizes (LDS
//                           histograms, one global atomic per counter and
//                           workgroup)
//   msm_part_scan_kernel    partition offsets over all windows, and the
//                           partitions' slices (runs of at most 2^15 of their
//                           pairs, the unit of work of the next passes)
//   msm_partition_kernel    per 4096-pair chunk: local sort by partition,
//                           one atomic per partition reserves the chunk's run
//   msm_bucket_count_kernel per slice: bucket counts (LDS histogram, one
//                           global atomic per bucket and slice)
//   msm_bucket_offsets_kernel per partition: bucket offsets and cursors, and
//                           each bucket's segments (runs of its entries, the
//                           bucket sums' unit of work)
//   msm_bucket_scatter_kernel per slice: reserves each bucket's run with one
//                           atomic, then writes the entries i | sign << 31 in
//                           bucket order
// A partition that holds most of a window's pairs (skewed scalars) is so
// spread over many workgroups; the order of a bucket's entries is that of the
// reservations, which the sums do not depend on.
constexpr uint32_t MSM_KEY_NONE = 0xFFFFu;
constexpr uint32_t MSM_MAX_PARTS = 128;           // 2^(15 - 8): c <= 16
constexpr uint32_t MSM_MAX_COUNTERS = 2048;       // windows x partitions, any c in 4..16
constexpr uint32_t SORT_THREADS = 256;
constexpr uint32_t SORT_PER_THREAD = 16;
constexpr uint32_t SORT_CHUNK = SORT_THREADS * SORT_PER_THREAD;
constexpr uint32_t SORT_SLICE_LOG = 15;  // pairs of one partition per count / scatter workgroup
// A bucket segment is a run of entries of one bucket, the bucket sums' unit
// of work: a bucket of more entries (skewed scalars, a sparse top window) is
// summed by several threads and their partial sums merged. A plan's segment
// lengths come packed in one word, seg_code (msm_seg_code): bits 0-7 the log2
// of an ordinary segment, four times the mean bucket size so that a bucket of
// uniformly random digits is one segment; bits 8-15 the log2 of a heavy
// bucket's segments (a heavy bucket has more than two ordinary segments):
// short where few points can make a bucket heavy, so that its threads' chains
// of additions stay short, and longer for the largest MSMs, whose heavy
// buckets (all-equal scalars) would otherwise make millions of segments.
__device__ __forceinline__ uint32_t bucket_seg_log(uint32_t count, uint32_t seg_code) {
  const uint32_t lg = seg_code & 0xFFu;
  return count > (2u << lg) ? (seg_code >> 8) & 0xFFu : lg;
}
__device__ __forceinline__ uint32_t bucket_segs(uint32_t count, uint32_t seg_code) {
  const uint32_t lg = bucket_seg_log(count, seg_code);
  return (count + (1u << lg) - 1) >> lg;
}

// Per point: base in limb form; signed digits of the scalar as keys
// (keys[w * n + i]); partition sizes part_counts[w * parts + hi] (zeroed by
// the caller). A grid-stride loop, so each workgroup's LDS histogram covers
// many points.
__global__ void __launch_bounds__(256) msm_digits_kernel(const uint4* __restrict__ scalars,
                                                         const uint4* __restrict__ bases, uint32_t n, uint32_t c,
                                                         uint32_t windows, uint32_t lo_bits, BaseLimbs* __restrict__ blimbs,
                                                         uint16_t* __restrict__ keys, uint32_t* __restrict__ part_counts) {
  __shared__ uint32_t hist[MSM_MAX_COUNTERS];
  const uint32_t parts = 1u << (c - 1 - lo_bits);
  const uint32_t nctr = windows * parts;
  for (uint32_t k = threadIdx.x; k < nctr; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  const uint32_t half = 1u << (c - 1);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint4 bx0 = bases[4 * (size_t)i], bx1 = bases[4 * (size_t)i + 1];
    const uint4 by0 = bases[4 * (size_t)i + 2], by1 = bases[4 * (size_t)i + 3];
    const uint32_t xw[8] = {bx0.x, bx0.y, bx0.z, bx0.w, bx1.x, bx1.y, bx1.z, bx1.w};
    const uint32_t yw[8] = {by0.x, by0.y, by0.z, by0.w, by1.x, by1.y, by1.z, by1.w};
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) any |= xw[k] | yw[k];
    const bool inf = any == 0;  // halo2's affine identity: x = y = 0
    uint32_t xc[8], yc[8];
    fq::q_canonical_words(xc, fq::q_from_words_shift(xw));
    fq::q_canonical_words(yc, fq::q_from_words_shift(yw));
    BaseLimbs e;
    e.x0 = make_uint4(xc[0], xc[1], xc[2], xc[3]);
    e.x1 = make_uint4(xc[4], xc[5], xc[6], xc[7]);
    e.y0 = make_uint4(yc[0], yc[1], yc[2], yc[3]);
    e.y1 = make_uint4(yc[4], yc[5], yc[6], yc[7] | (inf ? 0x80000000u : 0u));
    blimbs[i] = e;
    // canonical scalar
    const uint4 s0 = scalars[2 * (size_t)i], s1 = scalars[2 * (size_t)i + 1];
    const uint32_t sw[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    uint32_t cw[8];
    to_canonical_words(cw, mont_mul(unpack_words(sw), fr_from_mont_factor()));
    uint32_t carry = 0;
    for (uint32_t w = 0; w < windows; ++w) {
      const uint32_t v = scalar_bits(cw, w * c, c) + carry;
      // signed digit: above half it becomes v - 2^c with a carry into the next window
      const bool neg = v > half;
      const uint32_t mag = neg ? (1u << c) - v : v;
      carry = neg ? 1u : 0u;
      uint32_t key = MSM_KEY_NONE;
      if (mag != 0 && !inf) {
        key = (mag - 1u) | (neg ? 0x8000u : 0u);
        atomicAdd(&hist[w * parts + ((mag - 1u) >> lo_bits)], 1u);
      }
      keys[(size_t)w * n + i] = (uint16_t)key;
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nctr; k += blockDim.x)
    if (hist[k]) atomicAdd(&part_counts[k], hist[k]);
}

// Exclusive scan of the m <= MSM_MAX_COUNTERS partition sizes (window-major)
// into part_off[0..m] (part_off[m] = the total), and the same into
// part_cursor (the partition kernel's reservation cursors; the two may be
// one array). With slice_off, also the exclusive scan of the partitions'
// slice counts ceil(size / 2^SORT_SLICE_LOG) into slice_off[0..m]. One
// workgroup of 256. Also scans the partitions' segment counts.
__global__ void __launch_bounds__(256) msm_part_scan_kernel(const uint32_t* __restrict__ part_counts, uint32_t m,
                                                            uint32_t* part_off, uint32_t* part_cursor,
                                                            uint32_t* __restrict__ slice_off) {
  __shared__ uint32_t part[256], spart[256];
  const uint32_t per = (m + 255) / 256;
  const uint32_t lo = threadIdx.x * per, hi = min(m, lo + per);
  auto slices = [](uint32_t v) { return (v + (1u << SORT_SLICE_LOG) - 1) >> SORT_SLICE_LOG; };
  uint32_t sum = 0, ssum = 0;
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t v = part_counts[k];
    sum += v;
    ssum += slices(v);
  }
  part[threadIdx.x] = sum;
  spart[threadIdx.x] = ssum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0, srun = 0;
    for (uint32_t t = 0; t < 256; ++t) {
      const uint32_t v = part[t], sv = spart[t];
      part[t] = run;
      spart[t] = srun;
      run += v;
      srun += sv;
    }
  }
  __syncthreads();
  uint32_t run = part[threadIdx.x], srun = spart[threadIdx.x];
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t v = part_counts[k];
    part_off[k] = run;
    part_cursor[k] = run;
    run += v;
    if (slice_off) {
      slice_off[k] = srun;
      srun += slices(v);
    }
  }
  if (threadIdx.x == 255) {  // its range ends at m: run is the total
    part_off[m] = run;
    part_cursor[m] = run;
    if (slice_off) slice_off[m] = srun;
  }
}

// One workgroup per (chunk of SORT_CHUNK points, window): the chunk's pairs
// sorted by partition in LDS, then written to the partitions' runs reserved
// with one atomic each: t_entry = i | sign << 31, t_lo = the bucket's low bits.
__global__ void __launch_bounds__(SORT_THREADS) msm_partition_kernel(const uint16_t* __restrict__ keys, uint32_t n,
                                                                     uint32_t c, uint32_t lo_bits,
                                                                     uint32_t* __restrict__ part_cursor,
                                                                     uint32_t* __restrict__ t_entry,
                                                                     uint8_t* __restrict__ t_lo) {
  __shared__ uint32_t cnt[MSM_MAX_PARTS], off[MSM_MAX_PARTS], base[MSM_MAX_PARTS];
  __shared__ uint32_t st_entry[SORT_CHUNK];
  __shared__ uint8_t st_lo[SORT_CHUNK], st_part[SORT_CHUNK];
  const uint32_t w = blockIdx.y, i0 = blockIdx.x * SORT_CHUNK;
  const uint32_t parts = 1u << (c - 1 - lo_bits), lo_mask = (1u << lo_bits) - 1u;
  for (uint32_t p = threadIdx.x; p < parts; p += SORT_THREADS) cnt[p] = 0;
  __syncthreads();
  uint32_t key[SORT_PER_THREAD], rank[SORT_PER_THREAD];
#pragma unroll
  for (uint32_t q = 0; q < SORT_PER_THREAD; ++q) {
    const uint32_t i = i0 + q * SORT_THREADS + threadIdx.x;
    key[q] = i < n ? keys[(size_t)w * n + i] : MSM_KEY_NONE;
    rank[q] = key[q] != MSM_KEY_NONE ? atomicAdd(&cnt[(key[q] & 0x7fffu) >> lo_bits], 1u) : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t p = 0; p < parts; ++p) {
      off[p] = run;
      run += cnt[p];
    }
  }
  for (uint32_t p = threadIdx.x; p < parts; p += SORT_THREADS)
    base[p] = cnt[p] ? atomicAdd(&part_cursor[w * parts + p], cnt[p]) : 0u;
  __syncthreads();
#pragma unroll
  for (uint32_t q = 0; q < SORT_PER_THREAD; ++q) {
    if (key[q] == MSM_KEY_NONE) continue;
    const uint32_t b = key[q] & 0x7fffu, p = b >> lo_bits;
    const uint32_t j = off[p] + rank[q];
    st_entry[j] = (i0 + q * SORT_THREADS + threadIdx.x) | ((key[q] & 0x8000u) << 16);
    st_lo[j] = (uint8_t)(b & lo_mask);
    st_part[j] = (uint8_t)p;
  }
  __syncthreads();
  const uint32_t total = off[parts - 1] + cnt[parts - 1];
  for (uint32_t j = threadIdx.x; j < total; j += SORT_THREADS) {
    const uint32_t p = st_part[j];
    const uint32_t pos = base[p] + (j - off[p]);
    t_entry[pos] = st_entry[j];
    t_lo[pos] = st_lo[j];
  }
}

// The slice a count / scatter workgroup covers: partition g (the largest
// with slice_off[g] <= the workgroup; empty partitions have no slices) and
// the pairs [*start, *end) of the partitioned arrays. False past the last
// slice (the grid is a bound on the slices).
__device__ __forceinline__ bool sort_slice(const uint32_t* slice_off, const uint32_t* part_off,
                                           const uint32_t* part_counts, uint32_t m, uint32_t wg, uint32_t* g,
                                           uint32_t* start, uint32_t* end) {
  if (wg >= slice_off[m]) return false;
  uint32_t lo = 0, hi = m;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (slice_off[mid] <= wg) lo = mid; else hi = mid;
  }
  *g = lo;
  const uint32_t first = part_off[lo] + ((wg - slice_off[lo]) << SORT_SLICE_LOG);
  *start = first;
  *end = min(part_off[lo] + part_counts[lo], first + (1u << SORT_SLICE_LOG));
  return true;
}

// One workgroup per slice: the slice's bucket counts in LDS, kept for the
// scatter (slice_hist, 2^lo_bits words per slice) and added to the global
// counts (bucket g * 2^lo_bits + low bits for partition g; zeroed by the
// caller).
__global__ void __launch_bounds__(SORT_THREADS) msm_bucket_count_kernel(
    const uint32_t* __restrict__ slice_off, const uint32_t* __restrict__ part_off,
    const uint32_t* __restrict__ part_counts, uint32_t m, const uint8_t* __restrict__ t_lo, uint32_t lo_bits,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ slice_hist) {
  __shared__ uint32_t hist[256];
  uint32_t g, start, end;
  if (!sort_slice(slice_off, part_off, part_counts, m, blockIdx.x, &g, &start, &end)) return;
  const uint32_t nlo = 1u << lo_bits;
  for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) hist[l] = 0;
  __syncthreads();
  for (uint32_t j = start + threadIdx.x; j < end; j += SORT_THREADS) atomicAdd(&hist[t_lo[j]], 1u);
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) {
    slice_hist[(blockIdx.x << lo_bits) + l] = hist[l];
    if (hist[l]) atomicAdd(&counts[(g << lo_bits) + l], hist[l]);
  }
}

// One workgroup per partition: its buckets' list offsets (and the scatter's
// cursors, the same), and their segments: each bucket's first segment within
// the partition, the partition's segment count, and the device maximum of
// the segments of one bucket.
__global__ void __launch_bounds__(SORT_THREADS) msm_bucket_offsets_kernel(
    const uint32_t* __restrict__ part_off, const uint32_t* __restrict__ counts, uint32_t lo_bits, uint32_t seg_code,
    uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor, uint32_t* __restrict__ seg_local,
    uint32_t* __restrict__ part_segs, uint32_t* __restrict__ seg_max) {
  __shared__ uint32_t lcnt[256], cur[256], loff[256];
  const uint32_t g = blockIdx.x, nlo = 1u << lo_bits, b0 = g << lo_bits;
  for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) lcnt[l] = counts[b0 + l];
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = part_off[g], segs = 0, most = 0;
    for (uint32_t l = 0; l < nlo; ++l) {
      cur[l] = run;
      run += lcnt[l];
      loff[l] = segs;  // the bucket's first segment within the partition
      const uint32_t bs = bucket_segs(lcnt[l], seg_code);
      segs += bs;
      most = max(most, bs);
    }
    part_segs[g] = segs;
    if (most > 1) atomicMax(seg_max, most);
  }
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) {
    offsets[b0 + l] = cur[l];
    cursor[b0 + l] = cur[l];
    seg_local[b0 + l] = loff[l];
  }
}

// One workgroup per slice: with the slice's bucket counts (slice_hist), one
// atomic per bucket on the cursors reserves the slice's run of each bucket,
// then the slice goes
// into those runs a SORT_CHUNK chunk at a time, sorted by bucket in LDS so
// that consecutive threads write each run.
__global__ void __launch_bounds__(SORT_THREADS) msm_bucket_scatter_kernel(
    const uint32_t* __restrict__ slice_off, const uint32_t* __restrict__ part_off,
    const uint32_t* __restrict__ part_counts, uint32_t m, const uint32_t* __restrict__ t_entry,
    const uint8_t* __restrict__ t_lo, uint32_t lo_bits, const uint32_t* __restrict__ slice_hist,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ list) {
  __shared__ uint32_t cur[256], lcnt[256], loff[256];
  __shared__ uint32_t st_entry[SORT_CHUNK];
  __shared__ uint8_t st_lo[SORT_CHUNK];
  uint32_t g, start, end;
  if (!sort_slice(slice_off, part_off, part_counts, m, blockIdx.x, &g, &start, &end)) return;
  const uint32_t nlo = 1u << lo_bits;
  // the slice's bucket counts, from the count kernel
  for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) {
    const uint32_t h = slice_hist[(blockIdx.x << lo_bits) + l];
    cur[l] = h ? atomicAdd(&cursor[(g << lo_bits) + l], h) : 0u;
  }
  for (uint32_t c0 = start; c0 < end; c0 += SORT_CHUNK) {
    const uint32_t len = min(SORT_CHUNK, end - c0);
    __syncthreads();  // cur is set; the previous chunk's reads of lcnt / loff / staging are done
    for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) lcnt[l] = 0;
    __syncthreads();
    uint32_t lo[SORT_PER_THREAD], rank[SORT_PER_THREAD], ent[SORT_PER_THREAD];
#pragma unroll
    for (uint32_t q = 0; q < SORT_PER_THREAD; ++q) {
      const uint32_t j = q * SORT_THREADS + threadIdx.x;
      if (j < len) {
        lo[q] = t_lo[c0 + j];
        ent[q] = t_entry[c0 + j];
        rank[q] = atomicAdd(&lcnt[lo[q]], 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (uint32_t l = 0; l < nlo; ++l) {
        loff[l] = run;
        run += lcnt[l];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < SORT_PER_THREAD; ++q) {
      const uint32_t j = q * SORT_THREADS + threadIdx.x;
      if (j < len) {
        st_entry[loff[lo[q]] + rank[q]] = ent[q];
        st_lo[loff[lo[q]] + rank[q]] = (uint8_t)lo[q];
      }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < len; j += SORT_THREADS) {
      const uint32_t l = st_lo[j];
      list[cur[l] + (j - loff[l])] = st_entry[j];
    }
    __syncthreads();
    // advance each bucket's cursor past this chunk's entries
    for (uint32_t l = threadIdx.x; l < nlo; l += SORT_THREADS) cur[l] += lcnt[l];
  }
}

// Global index of bucket b's first segment: the segment offset of its
// partition (b >> lo_bits) plus its offset inside the partition.
__device__ __forceinline__ uint32_t seg_prefix(const uint32_t* part_seg_off, const uint32_t* seg_local, uint32_t b,
                                               uint32_t lo_bits) {
  return part_seg_off[b >> lo_bits] + seg_local[b];
}

// The bucket of global segment sidx: the largest b whose first segment is
// <= sidx (empty buckets have no segments, so that bucket is never empty).
__device__ __forceinline__ uint32_t segment_bucket(const uint32_t* part_seg_off, const uint32_t* seg_local,
                                                   uint32_t lo_bits, uint32_t nbuckets, uint32_t sidx) {
  uint32_t lo = 0, hi = nbuckets;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (seg_prefix(part_seg_off, seg_local, mid, lo_bits) <= sidx) lo = mid; else hi = mid;
  }
  return lo;
}

// Segments are handed to the bucket-sum threads longest first, so that the
// 64 lanes of a wave sum segments of about the same length (a wave runs as
// long as its longest lane; with random digits the bucket sizes of one wave
// spread over +-2.4 standard deviations) and the longest waves start first.
// The order is a counting sort on SEG_BINS bins of the segment length
// (shifted so that the longest ordinary segment, 2^(lg + 1) entries, fits).
constexpr uint32_t SEG_BINS = 1024;  // > 2^(lg + 1) for lg < 10; longer segments binned coarser
__device__ __forceinline__ uint32_t segment_bin(uint32_t len, uint32_t seg_code) {
  const uint32_t lg = seg_code & 0xFFu;
  const uint32_t shift = lg + 1 > 10 ? lg + 1 - 10 : 0u;
  return SEG_BINS - 1 - min(len >> shift, SEG_BINS - 1);  // longest first
}
__device__ __forceinline__ uint32_t segment_len(const uint32_t* part_seg_off, const uint32_t* seg_local,
                                                const uint32_t* counts, uint32_t lo_bits, uint32_t seg_code,
                                                uint32_t b, uint32_t sidx) {
  const uint32_t count = counts[b], lg = bucket_seg_log(count, seg_code);
  const uint32_t first = (sidx - seg_prefix(part_seg_off, seg_local, b, lo_bits)) << lg;
  return min(1u << lg, count - first);
}

// One thread per segment: the per-bin segment counts (LDS histogram, one
// global atomic per bin and workgroup; seg_hist zeroed by the caller).
__global__ void __launch_bounds__(256) msm_seg_hist_kernel(const uint32_t* __restrict__ part_seg_off,
                                                           const uint32_t* __restrict__ seg_local,
                                                           const uint32_t* __restrict__ counts, uint32_t nparts,
                                                           uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                                                           uint32_t* __restrict__ seg_hist) {
  __shared__ uint32_t h[SEG_BINS];
  const uint32_t total = part_seg_off[nparts];
  if (blockIdx.x * blockDim.x >= total) return;
  for (uint32_t i = threadIdx.x; i < SEG_BINS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx < total) {
    const uint32_t b = segment_bucket(part_seg_off, seg_local, lo_bits, nbuckets, sidx);
    atomicAdd(&h[segment_bin(segment_len(part_seg_off, seg_local, counts, lo_bits, seg_code, b, sidx), seg_code)],
              1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < SEG_BINS; i += blockDim.x)
    if (h[i]) atomicAdd(&seg_hist[i], h[i]);
}

// One thread per segment: its place in the order (bin offsets from the
// scan of seg_hist, as cursors: one atomic per bin and workgroup reserves
// the workgroup's places). Writes the order as (segment, bucket) pairs.
__global__ void __launch_bounds__(256) msm_seg_order_kernel(const uint32_t* __restrict__ part_seg_off,
                                                            const uint32_t* __restrict__ seg_local,
                                                            const uint32_t* __restrict__ counts, uint32_t nparts,
                                                            uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                                                            uint32_t* __restrict__ bin_cursor,
                                                            uint32_t* __restrict__ order_seg,
                                                            uint32_t* __restrict__ order_bucket) {
  __shared__ uint32_t h[SEG_BINS], base[SEG_BINS];
  const uint32_t total = part_seg_off[nparts];
  if (blockIdx.x * blockDim.x >= total) return;
  for (uint32_t i = threadIdx.x; i < SEG_BINS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = sidx < total;
  uint32_t b = 0, bin = 0, rank = 0;
  if (valid) {
    b = segment_bucket(part_seg_off, seg_local, lo_bits, nbuckets, sidx);
    bin = segment_bin(segment_len(part_seg_off, seg_local, counts, lo_bits, seg_code, b, sidx), seg_code);
    rank = atomicAdd(&h[bin], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < SEG_BINS; i += blockDim.x) base[i] = h[i] ? atomicAdd(&bin_cursor[i], h[i]) : 0u;
  __syncthreads();
  if (valid) {
    const uint32_t pos = base[bin] + rank;
    order_seg[pos] = sidx;
    order_bucket[pos] = b;
  }
}

// One thread per bucket segment (a run of one bucket's entries), in the
// order of msm_seg_order_kernel: the sum of its points, accumulated in XYZZ
// coordinates and stored so (in the bucket array for a bucket of one
// segment, else at the segment's index).
__global__ void __launch_bounds__(64) msm_buckets_kernel(const BaseLimbs* __restrict__ blimbs, const uint32_t* __restrict__ offsets,
                                   const uint32_t* __restrict__ counts, const uint32_t* __restrict__ list,
                                   const uint32_t* __restrict__ part_seg_off, const uint32_t* __restrict__ seg_local,
                                   uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                                   const uint32_t* __restrict__ order_seg, const uint32_t* __restrict__ order_bucket,
                                   PointLimbs* __restrict__ seg_sums, uint32_t* __restrict__ seg_bucket,
                                   PointLimbs* __restrict__ buckets) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= part_seg_off[nparts]) return;
  const uint32_t sidx = order_seg[t], b = order_bucket[t];
  seg_bucket[sidx] = b;
  const uint32_t count = counts[b], lg = bucket_seg_log(count, seg_code);
  const uint32_t first = (sidx - seg_prefix(part_seg_off, seg_local, b, lo_bits)) << lg;
  const uint32_t start = offsets[b] + first, cnt = min(1u << lg, count - first);
  g1xyzz acc;
  bool inf = true;
  // software pipeline: the list entry two ahead and the base record one
  // ahead are in flight while the current addition runs. The record stays
  // packed until its iteration, and the loads are unconditional (indices
  // clamped to the segment's last entry): unpacking a record where it is
  // loaded, or a load under a branch, would make the wave wait for the load
  // right there. A segment has at least one entry.
  const uint32_t last = start + cnt - 1;
  uint32_t e = list[start];
  uint32_t e_next = list[min(start + 1, last)];
  BaseLimbs raw = blimbs[e & 0x7fffffffu];
  for (uint32_t k = 0; k < cnt; ++k) {
    const g1a cur = unpack_base(raw);
    const bool neg = (e >> 31) != 0;
    raw = blimbs[e_next & 0x7fffffffu];
    e = e_next;
    e_next = list[min(start + k + 2, last)];
    g1_bucket_add(acc, inf, cur, neg);
  }
  // a bucket of one segment (nearly every bucket with random scalars) is
  // complete: its sum goes to the bucket array; the others' segment sums are
  // merged by the tree and copied by the gather
  const g1xyzz sum = inf ? g1_xyzz_identity() : acc;
  if (cnt == count) store_point(buckets, b, sum);
  else store_point(seg_sums, sidx, sum);
}

// One level of the pairwise reduction of each bucket's segment sums: the
// segment at local index i (a multiple of 2 x 2^level) adds the one at
// i + 2^level, if the bucket has it. After the levels 0..L with 2^(L+1) >=
// the most segments of any bucket, a bucket's first segment holds its sum.
// A level whose stride 2^level is at least seg_max (the most segments of one
// bucket, from the sort) exits at once: the host launches the levels a bucket
// of n entries could need. (Eight to one per level took fewer launches but
// lengthened the chains of the active levels: slower with skewed scalars.)
__global__ void __launch_bounds__(256) msm_segment_tree_kernel(
    PointLimbs* __restrict__ seg_sums, const uint32_t* __restrict__ seg_bucket, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ part_seg_off, const uint32_t* __restrict__ seg_local,
    const uint32_t* __restrict__ seg_max, uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t level) {
  const uint32_t step = 1u << level;
  if (*seg_max <= step) return;
  const uint32_t total = part_seg_off[nparts];
  for (uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x; sidx < total; sidx += gridDim.x * blockDim.x) {
    const uint32_t b = seg_bucket[sidx];
    const uint32_t i = sidx - seg_prefix(part_seg_off, seg_local, b, lo_bits);
    if ((i & (2 * step - 1)) != 0) continue;
    if (i + step >= bucket_segs(counts[b], seg_code)) continue;
    store_point(seg_sums, sidx, g1_xyzz_add(load_point(seg_sums, sidx), load_point(seg_sums, sidx + step)));
  }
}

// One thread per bucket: the sum of a bucket of several segments (its first
// segment's, after the reduction) or the identity for an empty bucket; a
// bucket of one segment has its sum from the bucket kernel.
__global__ void msm_bucket_gather_kernel(const PointLimbs* __restrict__ seg_sums, const uint32_t* __restrict__ counts,
                                         const uint32_t* __restrict__ part_seg_off,
                                         const uint32_t* __restrict__ seg_local, uint32_t lo_bits, uint32_t seg_code,
                                         uint32_t nbuckets, PointLimbs* __restrict__ buckets) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbuckets) return;
  const uint32_t segs = bucket_segs(counts[b], seg_code);
  if (segs == 0) store_point(buckets, b, g1_xyzz_identity());
  else if (segs > 1) store_point(buckets, b, load_point(seg_sums, seg_prefix(part_seg_off, seg_local, b, lo_bits)));
  // one segment: the bucket kernel stored the sum
}

// Two threads, in different waves, per run of `run` buckets [lo, lo + run)
// of window w: sum_{j in run} (j + 1) B_j = T + lo * S, where the running
// sums over the run (from its top down) give T = sum (j - lo + 1) B_j, and
// S = sum B_j. One thread computes T (two additions per bucket), the other S
// and lo * S (one addition per bucket, then double-and-add), so that neither
// chain is longer than about two additions per bucket: the kernel has few
// threads, and its time is the chain's latency. Outputs T and lo * S as
// parts 2r and 2r + 1 of run r. Workgroups of 64: workgroup 2q + role takes
// runs 64q .. 64q + 63.
__global__ void __launch_bounds__(64) msm_chunks_kernel(const PointLimbs* __restrict__ buckets, uint32_t half,
                                                        uint32_t run, uint32_t runs_per_window, uint32_t total_runs,
                                                        PointLimbs* __restrict__ out) {
  const uint32_t role = blockIdx.x & 1u, r = (blockIdx.x >> 1) * 64 + threadIdx.x;
  if (r >= total_runs) return;
  const uint32_t w = r / runs_per_window, k = r % runs_per_window;
  const uint32_t lo = k * run;
  g1xyzz S = g1_xyzz_identity();
  if (role == 0) {
    g1xyzz T = g1_xyzz_identity();
    for (uint32_t j = run; j-- > 0;) {
      S = g1_xyzz_add(S, load_point(buckets, w * half + lo + j));
      T = g1_xyzz_add(T, S);
    }
    store_point(out, 2 * r, T);
  } else {
    for (uint32_t j = run; j-- > 0;) S = g1_xyzz_add(S, load_point(buckets, w * half + lo + j));
    // lo * S by double-and-add, most significant bit first (doubling the
    // identity, ZZ = 0, gives ZZ = 0 again)
    g1xyzz L = lo ? S : g1_xyzz_identity();
    for (int bit = lo ? 30 - __clz(lo) : -1; bit >= 0; --bit) {
      g1_xyzz_double(L);
      if ((lo >> bit) & 1u) L = g1_xyzz_add(L, S);
    }
    store_point(out, 2 * r + 1, L);
  }
}

// One workgroup (512 threads) per window: tree sum of its parts (two per
// run); the window's sum goes out as canonical halo2 words X, Y, Z (96 bytes).
constexpr uint32_t WINDOW_THREADS = 512;
__global__ void __launch_bounds__(WINDOW_THREADS) msm_windows_kernel(const PointLimbs* __restrict__ parts,
                                                                     uint32_t parts_per_window, uint32_t count,
                                                                     uint32_t stride, uint4* __restrict__ out) {
  __shared__ PointLimbs tree[WINDOW_THREADS];
  const uint32_t w = blockIdx.x;
  g1xyzz acc = g1_xyzz_identity();
  for (uint32_t k = threadIdx.x; k < count; k += WINDOW_THREADS)
    acc = g1_xyzz_add(acc, load_point(parts, w * parts_per_window + k * stride));
  store_point(tree, threadIdx.x, acc);
  __syncthreads();
  for (uint32_t off = WINDOW_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const g1xyzz a = g1_xyzz_add(load_point(tree, threadIdx.x), load_point(tree, threadIdx.x + off));
      store_point(tree, threadIdx.x, a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // Jacobian (X ZZ^2, Y ZZZ^2, ZZZ); the identity (ZZ = ZZZ = 0) maps to Z = 0
    const g1j s = g1_xyzz_to_jacobian(load_point(tree, 0));
    uint32_t x[8], y[8], z[8];
    fq::q_to_words(x, s.x);
    fq::q_to_words(y, s.y);
    fq::q_to_words(z, s.z);
    uint4* o = out + 6 * w;
    o[0] = make_uint4(x[0], x[1], x[2], x[3]);
    o[1] = make_uint4(x[4], x[5], x[6], x[7]);
    o[2] = make_uint4(y[0], y[1], y[2], y[3]);
    o[3] = make_uint4(y[4], y[5], y[6], y[7]);
    o[4] = make_uint4(z[0], z[1], z[2], z[3]);
    o[5] = make_uint4(z[4], z[5], z[6], z[7]);
  }
}

// Windows of many parts (WINDOW_SLICE_MIN and up): each window's parts are
// first summed in slices by one-wave workgroups, spread over many CUs
// (the sum is a chain of dependent additions, so its time is its depth).
// Workgroup (w, g) sums slice g of window w (parts_per_window / slices
// entries from g times that) through its 64 lanes and a tree in LDS and
// leaves the sum in the slice's first entry (only this workgroup reads the
// slice, and every read is done before that store).
constexpr uint32_t WINDOW_SLICE_MIN = 4096;
constexpr uint32_t WINDOW_SLICES = 32;
__global__ void __launch_bounds__(64) msm_window_slices_kernel(PointLimbs* __restrict__ parts,
                                                                uint32_t parts_per_window) {
  __shared__ PointLimbs tree[64];
  const uint32_t w = blockIdx.x / WINDOW_SLICES, g = blockIdx.x % WINDOW_SLICES;
  const uint32_t len = parts_per_window / WINDOW_SLICES;
  PointLimbs* slice = parts + (size_t)w * parts_per_window + (size_t)g * len;
  g1xyzz acc = g1_xyzz_identity();
  for (uint32_t k = threadIdx.x; k < len; k += 64) acc = g1_xyzz_add(acc, load_point(slice, k));
  store_point(tree, threadIdx.x, acc);
  __syncthreads();
  for (uint32_t off = 32; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      acc = g1_xyzz_add(acc, load_point(tree, threadIdx.x + off));
      store_point(tree, threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) store_point(slice, 0, acc);
}

// Synthetic bases: point i = k_i G for the generator G = (1, 2), k_i a
// nonzero 64-bit integer from SplitMix64 of (seed, start + i); affine, halo2
// Fq Montgomery words (x then y).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void fill_random_points_kernel(uint4* __restrict__ out, uint64_t count, uint64_t seed, uint64_t start) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint64_t k = mix64(seed * 0xD1B54A32D192ED03ull + (start + t) * 0x9E3779B97F4A7C15ull) | 1ull;
  // G = (1, 2) in the 2^261 domain
  g1a g;
  g.x = fq::q_from_const(fq::ONE261);
  g.y = fq::q_dbl(g.x);
  g.inf = false;
  g1j acc = g1_identity();
  for (int bit = 63 - __clzll(k); bit >= 0; --bit) {
    acc = g1_double(acc);
    if ((k >> bit) & 1ull) acc = g1_add_mixed(acc, g);
  }
  // affine: z^-1 = z^(q-2)
  fe zi = fq::q_from_const(fq::ONE261);
  for (int w = 7; w >= 0; --w) {
    const uint32_t e = fq::p32(w) - (w == 0 ? 2u : 0u);
    for (int bit = 31; bit >= 0; --bit) {
      zi = fq::q_sqr(zi);
      if ((e >> bit) & 1u) zi = fq::q_mul(zi, acc.z);
    }
  }
  const fe zi2 = fq::q_sqr(zi), zi3 = fq::q_mul(zi2, zi);
  uint32_t x[8], y[8];
  fq::q_to_words(x, fq::q_mul(acc.x, zi2));
  fq::q_to_words(y, fq::q_mul(acc.y, zi3));
  uint4* o = out + 4 * t;
  o[0] = make_uint4(x[0], x[1], x[2], x[3]);
  o[1] = make_uint4(x[4], x[5], x[6], x[7]);
  o[2] = make_uint4(y[0], y[1], y[2], y[3]);
  o[3] = make_uint4(y[4], y[5], y[6], y[7]);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_fill_random_points(void* out, uint64_t count, uint64_t seed, uint64_t start, hipStream_t s) {
  fill_random_points_kernel<<<(uint32_t)((count + 63) / 64), 64, 0, s>>>((uint4*)out, count, seed, start);
  return hipGetLastError();
}

hipError_t launch_msm_digits(const void* scalars, const void* bases, uint32_t n, uint32_t c, uint32_t windows,
                             uint32_t lo_bits, void* blimbs, uint16_t* keys, uint32_t* part_counts, hipStream_t s) {
  if (c < 2 || c > 16 || lo_bits > 8 || lo_bits > c - 1 || windows * (1u << (c - 1 - lo_bits)) > MSM_MAX_COUNTERS)
    return hipErrorInvalidValue;  // the kernels' LDS arrays are sized for these bounds
  const uint32_t grid = std::min<uint32_t>((n + 255) / 256, 2048u);
  msm_digits_kernel<<<grid, 256, 0, s>>>((const uint4*)scalars, (const uint4*)bases, n, c, windows, lo_bits,
                                         (BaseLimbs*)blimbs, keys, part_counts);
  return hipGetLastError();
}
// upper bound of the slices of `pairs` pairs over m partitions
uint64_t msm_max_slices(uint64_t pairs, uint32_t m) { return (pairs >> SORT_SLICE_LOG) + m; }
hipError_t launch_msm_sort(const uint16_t* keys, uint32_t n, uint32_t c, uint32_t windows, uint32_t lo_bits,
                           const uint32_t* part_counts, uint32_t* part_off, uint32_t* part_cursor,
                           uint32_t* slice_off, uint32_t* t_entry, uint8_t* t_lo, uint32_t* counts, uint32_t* offsets,
                           uint32_t* cursor, uint32_t* list, uint32_t* slice_hist, uint32_t seg_code,
                           uint32_t* seg_local, uint32_t* part_segs, uint32_t* part_seg_off, uint32_t* seg_max,
                           hipStream_t s) {
  if (c < 2 || c > 16 || lo_bits > 8 || lo_bits > c - 1) return hipErrorInvalidValue;
  const uint32_t parts = 1u << (c - 1 - lo_bits), m = windows * parts;
  if (parts > MSM_MAX_PARTS || m > MSM_MAX_COUNTERS) return hipErrorInvalidValue;
  const uint32_t slices = (uint32_t)msm_max_slices((uint64_t)windows * n, m);
  msm_part_scan_kernel<<<1, 256, 0, s>>>(part_counts, m, part_off, part_cursor, slice_off);
  msm_partition_kernel<<<dim3((n + SORT_CHUNK - 1) / SORT_CHUNK, windows), SORT_THREADS, 0, s>>>(
      keys, n, c, lo_bits, part_cursor, t_entry, t_lo);
  msm_bucket_count_kernel<<<slices, SORT_THREADS, 0, s>>>(slice_off, part_off, part_counts, m, t_lo, lo_bits,
                                                          counts, slice_hist);
  msm_bucket_offsets_kernel<<<m, SORT_THREADS, 0, s>>>(part_off, counts, lo_bits, seg_code, offsets, cursor,
                                                       seg_local, part_segs, seg_max);
  msm_part_scan_kernel<<<1, 256, 0, s>>>(part_segs, m, part_seg_off, part_seg_off, nullptr);
  msm_bucket_scatter_kernel<<<slices, SORT_THREADS, 0, s>>>(slice_off, part_off, part_counts, m, t_entry, t_lo,
                                                            lo_bits, slice_hist, cursor, list);
  return hipGetLastError();
}
// Segment lengths (bucket_seg_log): ordinary 2^5..2^12 entries, four times
// the mean bucket size; heavy 2^4 entries below 2^18 (window, point) pairs,
// 2^6 up to 2^26 pairs, then longer, up to 2^8 from 2^28 pairs (never
// longer than the ordinary ones).
uint32_t msm_seg_code(uint64_t pairs, uint64_t nbuckets) {
  uint32_t lg = 5;
  while (lg < 12 && (1ull << lg) < 4 * pairs / nbuckets) ++lg;
  uint32_t heavy = pairs < (1ull << 18) ? 4 : 6;
  while (heavy < 8 && (pairs >> (20 + heavy)) > 1) ++heavy;
  return lg | std::min(heavy, lg) << 8;
}
// upper bound of the segments of `pairs` entries over `nbuckets` buckets
// (a bucket of c entries has at most c / 2^heavy + 1)
uint64_t msm_max_segments(uint64_t pairs, uint64_t nbuckets, uint32_t seg_code) {
  return (pairs >> ((seg_code >> 8) & 0xFFu)) + nbuckets;
}
uint32_t msm_lo_bits(uint32_t c) { return std::min(8u, c - 1); }
// The segments' order for the bucket sums (seg_hist, SEG_BINS words, zeroed
// by the caller; bin_cursor SEG_BINS + 1 words; order_*: max_segs words).
hipError_t launch_msm_seg_order(const uint32_t* part_seg_off, const uint32_t* seg_local, const uint32_t* counts,
                                uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                                uint64_t max_segs, uint32_t* seg_hist, uint32_t* bin_cursor, uint32_t* order_seg,
                                uint32_t* order_bucket, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((max_segs + 255) / 256);
  msm_seg_hist_kernel<<<blocks, 256, 0, s>>>(part_seg_off, seg_local, counts, nparts, lo_bits, seg_code, nbuckets,
                                             seg_hist);
  msm_part_scan_kernel<<<1, 256, 0, s>>>(seg_hist, SEG_BINS, bin_cursor, bin_cursor, nullptr);
  msm_seg_order_kernel<<<blocks, 256, 0, s>>>(part_seg_off, seg_local, counts, nparts, lo_bits, seg_code, nbuckets,
                                              bin_cursor, order_seg, order_bucket);
  return hipGetLastError();
}
uint32_t msm_seg_bins() { return SEG_BINS; }
hipError_t launch_msm_buckets(const void* blimbs, const uint32_t* offsets, const uint32_t* counts,
                              const uint32_t* list, const uint32_t* part_seg_off, const uint32_t* seg_local,
                              uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                              uint64_t max_segs, const uint32_t* order_seg, const uint32_t* order_bucket,
                              void* seg_sums, uint32_t* seg_bucket, void* buckets, hipStream_t s) {
  msm_buckets_kernel<<<(uint32_t)((max_segs + 63) / 64), 64, 0, s>>>(
      (const BaseLimbs*)blimbs, offsets, counts, list, part_seg_off, seg_local, nparts, lo_bits, seg_code, nbuckets,
      order_seg, order_bucket, (PointLimbs*)seg_sums, seg_bucket, (PointLimbs*)buckets);
  return hipGetLastError();
}
// The segment sums reduced to one sum per bucket: the tree levels a bucket
// of n entries could need (idle ones exit on seg_max), then the gather.
hipError_t launch_msm_bucket_merge(void* seg_sums, const uint32_t* seg_bucket, const uint32_t* counts,
                                   const uint32_t* part_seg_off, const uint32_t* seg_local, const uint32_t* seg_max,
                                   uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t n, uint32_t nbuckets,
                                   uint64_t max_segs, void* buckets, hipStream_t s) {
  // segments of a bucket of at most n entries: two, or n / 2^heavy for a heavy one
  const uint32_t heavy = (seg_code >> 8) & 0xFFu;
  const uint64_t most = std::max<uint64_t>(2, ((uint64_t)n + (1u << heavy) - 1) >> heavy);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(1024, (max_segs + 255) / 256);
  for (uint32_t level = 0; (1ull << level) < most; ++level)
    msm_segment_tree_kernel<<<blocks, 256, 0, s>>>((PointLimbs*)seg_sums, seg_bucket, counts, part_seg_off, seg_local,
                                                   seg_max, nparts, lo_bits, seg_code, level);
  msm_bucket_gather_kernel<<<(nbuckets + 63) / 64, 64, 0, s>>>((const PointLimbs*)seg_sums, counts, part_seg_off,
                                                               seg_local, lo_bits, seg_code, nbuckets,
                                                               (PointLimbs*)buckets);
  return hipGetLastError();
}
// parts: 2 x runs_per_window x windows point entries
hipError_t launch_msm_chunks(const void* buckets, uint32_t half, uint32_t run, uint32_t runs_per_window,
                             uint32_t windows, void* parts, hipStream_t s) {
  const uint32_t total = runs_per_window * windows;
  msm_chunks_kernel<<<2 * ((total + 63) / 64), 64, 0, s>>>((const PointLimbs*)buckets, half, run, runs_per_window,
                                                          total, (PointLimbs*)parts);
  return hipGetLastError();
}
// parts: overwritten where the windows are sliced (a slice's first entry
// takes the slice's sum)
hipError_t launch_msm_windows(void* parts, uint32_t runs_per_window, uint32_t windows, void* out, hipStream_t s) {
  const uint32_t per_window = 2 * runs_per_window;  // a power of two
  if (per_window >= WINDOW_SLICE_MIN) {
    msm_window_slices_kernel<<<windows * WINDOW_SLICES, 64, 0, s>>>((PointLimbs*)parts, per_window);
    msm_windows_kernel<<<windows, WINDOW_THREADS, 0, s>>>((const PointLimbs*)parts, per_window, WINDOW_SLICES,
                                                          per_window / WINDOW_SLICES, (uint4*)out);
  } else {
    msm_windows_kernel<<<windows, WINDOW_THREADS, 0, s>>>((const PointLimbs*)parts, per_window, per_window, 1,
                                                          (uint4*)out);
  }
  return hipGetLastError();
}
size_t msm_base_limb_bytes() { return sizeof(BaseLimbs); }
size_t msm_point_limb_bytes() { return sizeof(PointLimbs); }

}  // namespace zkgpu
