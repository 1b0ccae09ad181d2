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
//   msm_digits    one thread per point: base to limb form, scalar to signed
//                 digits, bucket histogram (one atomic per digit, whose old
//                 value is the point's rank in its bucket)
//   scan          exclusive prefix sum of the bucket counts (three launches)
//   msm_scatter   one thread per point: its index (sign in bit 31) into each
//                 bucket's list, at the bucket's offset plus its rank
//   msm_buckets   one thread per bucket: sum of its points (mixed additions)
//   msm_chunks    one thread per run of L buckets of a window: running sums
//                 give sum_j (j+1) B_j over the run
//   msm_windows   one workgroup per window: tree sum of its runs, stored as
//                 canonical halo2 words
//
// The windows are combined (Horner, c doublings each) by the host.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fq29.hpp"

namespace zkgpu {

// limb form of a base: x and y (2^261 domain) in 48-byte entries; bit 31 of
// the y entry's last word marks the identity
struct BaseLimbs {
  uint4 x0, x1, x2, y0, y1, y2;
};
// a bucket / partial sum: Jacobian X, Y, Z in 48-byte entries
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
__device__ __forceinline__ g1a load_base(const BaseLimbs* b, uint32_t i) {
  const BaseLimbs e = b[i];
  g1a p;
  p.x = fe_from3(e.x0, e.x1, e.x2);
  p.y = fe_from3(e.y0, e.y1, e.y2);
  p.inf = (e.y2.x >> 31) != 0;
  p.y.l[8] &= 0x7fffffffu;
  return p;
}
__device__ __forceinline__ g1j load_point(const PointLimbs* p, uint32_t i) {
  const PointLimbs e = p[i];
  return g1j{fe_from3(e.a[0], e.a[1], e.a[2]), fe_from3(e.a[3], e.a[4], e.a[5]), fe_from3(e.a[6], e.a[7], e.a[8])};
}
__device__ __forceinline__ void store_point(PointLimbs* p, uint32_t i, const g1j& v) {
  PointLimbs e;
  fe_to3(v.x, e.a[0], e.a[1], e.a[2]);
  fe_to3(v.y, e.a[3], e.a[4], e.a[5]);
  fe_to3(v.z, e.a[6], e.a[7], e.a[8]);
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

// Per point: base in limb form; signed digits of the scalar; bucket counts.
// digits[w * n + i] is the digit of window w (0 for identity bases),
// ranks[w * n + i] the point's rank among that bucket's points.
__global__ void msm_digits_kernel(const uint4* __restrict__ scalars, const uint4* __restrict__ bases, uint32_t n,
                                  uint32_t c, uint32_t windows, BaseLimbs* __restrict__ blimbs,
                                  int32_t* __restrict__ digits, uint32_t* __restrict__ ranks,
                                  uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 bx0 = bases[4 * (size_t)i], bx1 = bases[4 * (size_t)i + 1];
  const uint4 by0 = bases[4 * (size_t)i + 2], by1 = bases[4 * (size_t)i + 3];
  const uint32_t xw[8] = {bx0.x, bx0.y, bx0.z, bx0.w, bx1.x, bx1.y, bx1.z, bx1.w};
  const uint32_t yw[8] = {by0.x, by0.y, by0.z, by0.w, by1.x, by1.y, by1.z, by1.w};
  uint32_t any = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) any |= xw[k] | yw[k];
  const bool inf = any == 0;  // halo2's affine identity: x = y = 0
  BaseLimbs e;
  const fe x = fq::q_from_words(xw), y = fq::q_from_words(yw);
  fe_to3(x, e.x0, e.x1, e.x2);
  fe_to3(y, e.y0, e.y1, e.y2);
  if (inf) e.y2.x |= 0x80000000u;
  blimbs[i] = e;
  // canonical scalar
  const uint4 s0 = scalars[2 * (size_t)i], s1 = scalars[2 * (size_t)i + 1];
  const uint32_t sw[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
  uint32_t cw[8];
  to_canonical_words(cw, mont_mul(unpack_words(sw), fr_from_mont_factor()));
  const uint32_t half = 1u << (c - 1);
  uint32_t carry = 0;
  for (uint32_t w = 0; w < windows; ++w) {
    uint32_t v = scalar_bits(cw, w * c, c) + carry;
    int32_t d;
    if (v > half) {
      d = (int32_t)v - (int32_t)(1u << c);
      carry = 1;
    } else {
      d = (int32_t)v;
      carry = 0;
    }
    if (inf) d = 0;
    digits[(size_t)w * n + i] = d;
    if (d != 0) ranks[(size_t)w * n + i] = atomicAdd(&counts[w * half + (uint32_t)(d < 0 ? -d : d) - 1u], 1u);
  }
}

// Exclusive scan of counts[0..m): per block of 1024 entries its sum, then the
// block sums (one workgroup), then each block's entries.
constexpr uint32_t SCAN_BLOCK = 1024;
__global__ void scan_block_sums_kernel(const uint32_t* __restrict__ in, uint32_t m, uint32_t* __restrict__ sums) {
  __shared__ uint32_t part[256];
  const uint32_t base = blockIdx.x * SCAN_BLOCK;
  uint32_t s = 0;
  for (uint32_t k = threadIdx.x; k < SCAN_BLOCK; k += 256)
    if (base + k < m) s += in[base + k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}
// exclusive scan of sums[0..nb), nb <= 256 * 64, in one workgroup of 256
__global__ void scan_sums_kernel(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t part[256];
  const uint32_t per = (nb + 255) / 256;
  const uint32_t lo = threadIdx.x * per, hi = min(nb, lo + per);
  uint32_t s = 0;
  for (uint32_t k = lo; k < hi; ++k) s += sums[k];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t t = 0; t < 256; ++t) {
      const uint32_t v = part[t];
      part[t] = run;
      run += v;
    }
  }
  __syncthreads();
  uint32_t run = part[threadIdx.x];
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t v = sums[k];
    sums[k] = run;
    run += v;
  }
}
__global__ void scan_apply_kernel(const uint32_t* __restrict__ in, uint32_t m, const uint32_t* __restrict__ sums,
                                  uint32_t* __restrict__ out) {
  __shared__ uint32_t vals[SCAN_BLOCK];
  const uint32_t base = blockIdx.x * SCAN_BLOCK;
  for (uint32_t k = threadIdx.x; k < SCAN_BLOCK; k += 256) vals[k] = base + k < m ? in[base + k] : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = sums[blockIdx.x];
    for (uint32_t k = 0; k < SCAN_BLOCK; ++k) {
      const uint32_t v = vals[k];
      vals[k] = run;
      run += v;
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < SCAN_BLOCK; k += 256) {
    if (base + k < m) out[base + k] = vals[k];
  }
}

// Each point's index (bit 31: negative digit) into the list of every bucket
// it falls in.
__global__ void msm_scatter_kernel(const int32_t* __restrict__ digits, const uint32_t* __restrict__ ranks, uint32_t n,
                                   uint32_t c, uint32_t windows, const uint32_t* __restrict__ offsets,
                                   uint32_t* __restrict__ list) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t half = 1u << (c - 1);
  for (uint32_t w = 0; w < windows; ++w) {
    const int32_t d = digits[(size_t)w * n + i];
    if (d == 0) continue;
    const uint32_t key = w * half + (uint32_t)(d < 0 ? -d : d) - 1u;
    list[offsets[key] + ranks[(size_t)w * n + i]] = i | (d < 0 ? 0x80000000u : 0u);
  }
}

// One thread per bucket: the sum of its points.
__global__ void msm_buckets_kernel(const BaseLimbs* __restrict__ blimbs, const uint32_t* __restrict__ offsets,
                                   const uint32_t* __restrict__ counts, const uint32_t* __restrict__ list,
                                   uint32_t nbuckets, PointLimbs* __restrict__ buckets) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbuckets) return;
  const uint32_t start = offsets[b], cnt = counts[b];
  g1j acc = g1_identity();
  bool inf = true;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t e = list[start + k];
    g1_bucket_add(acc, inf, load_base(blimbs, e & 0x7fffffffu), (e >> 31) != 0);
  }
  if (inf) acc = g1_identity();
  store_point(buckets, b, acc);
}

// One thread per run of `run` buckets [lo, lo + run) of window w:
// sum_{j in run} (j + 1) B_j = T + lo * S, where the running sums over the run
// (from its top down) give S = sum B_j and T = sum (j - lo + 1) B_j.
__global__ void msm_chunks_kernel(const PointLimbs* __restrict__ buckets, uint32_t half, uint32_t run,
                                  uint32_t runs_per_window, uint32_t total_runs, PointLimbs* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total_runs) return;
  const uint32_t w = t / runs_per_window, k = t % runs_per_window;
  const uint32_t lo = k * run;
  g1j S = g1_identity(), T = g1_identity();
  for (uint32_t j = run; j-- > 0;) {
    S = g1_add(S, load_point(buckets, w * half + lo + j));
    T = g1_add(T, S);
  }
  // lo * S by double-and-add, most significant bit first
  g1j L = g1_identity();
  for (int bit = lo ? 31 - __clz(lo) : -1; bit >= 0; --bit) {
    L = g1_double(L);
    if ((lo >> bit) & 1u) L = g1_add(L, S);
  }
  store_point(out, t, g1_add(T, L));
}

// One workgroup (256 threads) per window: tree sum of its runs' partial sums;
// the window's sum goes out as canonical halo2 words X, Y, Z (96 bytes).
__global__ void __launch_bounds__(256) msm_windows_kernel(const PointLimbs* __restrict__ parts,
                                                          uint32_t runs_per_window, uint4* __restrict__ out) {
  __shared__ PointLimbs tree[256];
  const uint32_t w = blockIdx.x;
  g1j acc = g1_identity();
  for (uint32_t k = threadIdx.x; k < runs_per_window; k += 256)
    acc = g1_add(acc, load_point(parts, w * runs_per_window + k));
  store_point(tree, threadIdx.x, acc);
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const g1j a = g1_add(load_point(tree, threadIdx.x), load_point(tree, threadIdx.x + off));
      store_point(tree, threadIdx.x, a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const g1j s = load_point(tree, 0);
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
                             void* blimbs, int32_t* digits, uint32_t* ranks, uint32_t* counts, hipStream_t s) {
  msm_digits_kernel<<<(n + 255) / 256, 256, 0, s>>>((const uint4*)scalars, (const uint4*)bases, n, c, windows,
                                                    (BaseLimbs*)blimbs, digits, ranks, counts);
  return hipGetLastError();
}
hipError_t launch_msm_scan(const uint32_t* counts, uint32_t m, uint32_t* sums, uint32_t* offsets, hipStream_t s) {
  const uint32_t nb = (m + SCAN_BLOCK - 1) / SCAN_BLOCK;
  scan_block_sums_kernel<<<nb, 256, 0, s>>>(counts, m, sums);
  scan_sums_kernel<<<1, 256, 0, s>>>(sums, nb);
  scan_apply_kernel<<<nb, 256, 0, s>>>(counts, m, sums, offsets);
  return hipGetLastError();
}
hipError_t launch_msm_scatter(const int32_t* digits, const uint32_t* ranks, uint32_t n, uint32_t c, uint32_t windows,
                              const uint32_t* offsets, uint32_t* list, hipStream_t s) {
  msm_scatter_kernel<<<(n + 255) / 256, 256, 0, s>>>(digits, ranks, n, c, windows, offsets, list);
  return hipGetLastError();
}
hipError_t launch_msm_buckets(const void* blimbs, const uint32_t* offsets, const uint32_t* counts,
                              const uint32_t* list, uint32_t nbuckets, void* buckets, hipStream_t s) {
  msm_buckets_kernel<<<(nbuckets + 63) / 64, 64, 0, s>>>((const BaseLimbs*)blimbs, offsets, counts, list, nbuckets,
                                                        (PointLimbs*)buckets);
  return hipGetLastError();
}
hipError_t launch_msm_chunks(const void* buckets, uint32_t half, uint32_t run, uint32_t runs_per_window,
                             uint32_t windows, void* parts, hipStream_t s) {
  const uint32_t total = runs_per_window * windows;
  msm_chunks_kernel<<<(total + 63) / 64, 64, 0, s>>>((const PointLimbs*)buckets, half, run, runs_per_window, total,
                                                    (PointLimbs*)parts);
  return hipGetLastError();
}
hipError_t launch_msm_windows(const void* parts, uint32_t runs_per_window, uint32_t windows, void* out,
                              hipStream_t s) {
  msm_windows_kernel<<<windows, 256, 0, s>>>((const PointLimbs*)parts, runs_per_window, (uint4*)out);
  return hipGetLastError();
}
size_t msm_base_limb_bytes() { return sizeof(BaseLimbs); }
size_t msm_point_limb_bytes() { return sizeof(PointLimbs); }

}  // namespace zkgpu
