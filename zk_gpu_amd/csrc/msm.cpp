// G1 multi-scalar multiplication: plan, device workspace, launches, and the
// C-ABI entry points (parallel_msm_api of the reference, cuda/fft.cu:450-551,
// 578-582, and a device-resident variant).
//
// The device computes one sum per window (msm_kernels.hip); the host combines
// the W window sums by Horner's rule (c doublings and one addition per
// window, a few hundred point operations on 4 x 64-bit Montgomery words) and
// returns the result with Z = 1 (the identity as X = Y = Z = 0, as the
// reference leaves it).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/zkgpu_ntt.h"
#include "engine.hpp"

namespace zkgpu {

hipError_t launch_msm_digits(const void* scalars, const void* bases, uint32_t n, uint32_t c, uint32_t windows,
                             uint32_t lo_bits, void* blimbs, uint16_t* keys, uint32_t* part_counts, hipStream_t s);
hipError_t launch_msm_sort(const uint16_t* keys, uint32_t n, uint32_t c, uint32_t windows, uint32_t lo_bits,
                           const uint32_t* part_counts, uint32_t* part_off, uint32_t* part_cursor,
                           uint32_t* slice_off, uint32_t* t_entry, uint8_t* t_lo, uint32_t* counts, uint32_t* offsets,
                           uint32_t* cursor, uint32_t* list, uint32_t* slice_hist, uint32_t seg_code,
                           uint32_t* seg_local, uint32_t* part_segs, uint32_t* part_seg_off, uint32_t* seg_max,
                           hipStream_t s);
uint32_t msm_lo_bits(uint32_t c);
uint64_t msm_max_slices(uint64_t pairs, uint32_t m);
uint32_t msm_seg_code(uint64_t pairs, uint64_t nbuckets);
uint64_t msm_max_segments(uint64_t pairs, uint64_t nbuckets, uint32_t seg_code);
hipError_t launch_msm_seg_order(const uint32_t* part_seg_off, const uint32_t* seg_local, const uint32_t* counts,
                                uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                                uint64_t max_segs, uint32_t* seg_hist, uint32_t* bin_cursor, uint32_t* order_seg,
                                uint32_t* order_bucket, hipStream_t s);
uint32_t msm_seg_bins();
hipError_t launch_msm_buckets(const void* blimbs, const uint32_t* offsets, const uint32_t* counts,
                              const uint32_t* list, const uint32_t* part_seg_off, const uint32_t* seg_local,
                              uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t nbuckets,
                              uint64_t max_segs, const uint32_t* order_seg, const uint32_t* order_bucket,
                              void* seg_sums, uint32_t* seg_bucket, void* buckets, hipStream_t s);
hipError_t launch_msm_bucket_merge(void* seg_sums, const uint32_t* seg_bucket, const uint32_t* counts,
                                   const uint32_t* part_seg_off, const uint32_t* seg_local, const uint32_t* seg_max,
                                   uint32_t nparts, uint32_t lo_bits, uint32_t seg_code, uint32_t n, uint32_t nbuckets,
                                   uint64_t max_segs, void* buckets, hipStream_t s);
hipError_t launch_msm_chunks(const void* buckets, uint32_t half, uint32_t run, uint32_t runs_per_window,
                             uint32_t windows, void* parts, hipStream_t s);
hipError_t launch_msm_windows(void* parts, uint32_t runs_per_window, uint32_t windows, void* out, hipStream_t s);
size_t msm_base_limb_bytes();
size_t msm_point_limb_bytes();
hipError_t launch_fill_random_points(void* out, uint64_t count, uint64_t seed, uint64_t start, hipStream_t s);

namespace {

// ---------------------------------------------------------------- host Fq

// Fq elements as four little-endian u64 words in Montgomery form (R = 2^256),
// CIOS multiply over unsigned __int128.
struct Q {
  uint64_t w[4];
};
constexpr Q QMOD = {{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
constexpr uint64_t QINV = 0x87d20782e4866389ULL;  // -q^-1 mod 2^64
constexpr Q QZERO = {{0, 0, 0, 0}};
constexpr Q QONE = {{0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL}};

bool q_geq(const Q& a, const Q& b) {
  for (int i = 3; i >= 0; --i)
    if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
  return true;
}
void q_sub_raw(Q& a, const Q& b) {
  unsigned __int128 bw = 0;
  for (int i = 0; i < 4; ++i) {
    const unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - bw;
    a.w[i] = (uint64_t)d;
    bw = (d >> 64) & 1;
  }
}
Q q_add(const Q& a, const Q& b) {
  Q o;
  unsigned __int128 c = 0;
  for (int i = 0; i < 4; ++i) {
    const unsigned __int128 s = (unsigned __int128)a.w[i] + b.w[i] + c;
    o.w[i] = (uint64_t)s;
    c = s >> 64;
  }
  if (c || q_geq(o, QMOD)) q_sub_raw(o, QMOD);
  return o;
}
Q q_sub(const Q& a, const Q& b) {
  if (q_geq(a, b)) {
    Q o = a;
    q_sub_raw(o, b);
    return o;
  }
  Q o = a;  // a + q < 2^255: no carry out of the top word
  unsigned __int128 c = 0;
  for (int i = 0; i < 4; ++i) {
    const unsigned __int128 s = (unsigned __int128)o.w[i] + QMOD.w[i] + c;
    o.w[i] = (uint64_t)s;
    c = s >> 64;
  }
  q_sub_raw(o, b);
  return o;
}
Q q_mul(const Q& a, const Q& b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < 4; ++j) {
      const unsigned __int128 p = (unsigned __int128)a.w[i] * b.w[j] + t[j] + c;
      t[j] = (uint64_t)p;
      c = (uint64_t)(p >> 64);
    }
    unsigned __int128 s = (unsigned __int128)t[4] + c;
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    const uint64_t m = t[0] * QINV;
    unsigned __int128 p = (unsigned __int128)m * QMOD.w[0] + t[0];
    c = (uint64_t)(p >> 64);
    for (int j = 1; j < 4; ++j) {
      p = (unsigned __int128)m * QMOD.w[j] + t[j] + c;
      t[j - 1] = (uint64_t)p;
      c = (uint64_t)(p >> 64);
    }
    s = (unsigned __int128)t[4] + c;
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  Q o = {{t[0], t[1], t[2], t[3]}};
  if (t[4] || q_geq(o, QMOD)) q_sub_raw(o, QMOD);
  return o;
}
bool q_is_zero(const Q& a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0; }
bool q_eq(const Q& a, const Q& b) { return memcmp(a.w, b.w, 32) == 0; }
Q q_inv(const Q& a) {  // a^(q-2)
  Q e = QMOD;
  e.w[0] -= 2;
  Q acc = QONE;
  for (int i = 3; i >= 0; --i)
    for (int k = 63; k >= 0; --k) {
      acc = q_mul(acc, acc);
      if ((e.w[i] >> k) & 1) acc = q_mul(acc, a);
    }
  return acc;
}

struct Jac {
  Q x, y, z;
};
Jac j_identity() { return Jac{QZERO, QZERO, QZERO}; }
bool j_is_identity(const Jac& p) { return q_is_zero(p.z); }
Jac j_double(const Jac& p) {  // dbl-2009-l
  if (j_is_identity(p)) return p;
  const Q A = q_mul(p.x, p.x), B = q_mul(p.y, p.y), C = q_mul(B, B);
  const Q t = q_add(p.x, B);
  Q D = q_sub(q_sub(q_mul(t, t), A), C);
  D = q_add(D, D);
  const Q E = q_add(q_add(A, A), A);
  const Q F = q_mul(E, E);
  Jac o;
  o.x = q_sub(F, q_add(D, D));
  Q C8 = q_add(C, C);
  C8 = q_add(C8, C8);
  C8 = q_add(C8, C8);
  o.y = q_sub(q_mul(E, q_sub(D, o.x)), C8);
  const Q yz = q_mul(p.y, p.z);
  o.z = q_add(yz, yz);
  return o;
}
Jac j_add(const Jac& p, const Jac& q) {  // add-2007-bl
  if (j_is_identity(p)) return q;
  if (j_is_identity(q)) return p;
  const Q Z1Z1 = q_mul(p.z, p.z), Z2Z2 = q_mul(q.z, q.z);
  const Q U1 = q_mul(p.x, Z2Z2), U2 = q_mul(q.x, Z1Z1);
  const Q S1 = q_mul(p.y, q_mul(q.z, Z2Z2)), S2 = q_mul(q.y, q_mul(p.z, Z1Z1));
  const Q H = q_sub(U2, U1);
  Q R = q_sub(S2, S1);
  R = q_add(R, R);
  if (q_is_zero(H)) return q_is_zero(R) ? j_double(p) : j_identity();
  const Q H2 = q_add(H, H);
  const Q I = q_mul(H2, H2), J = q_mul(H, I), V = q_mul(U1, I);
  Jac o;
  o.x = q_sub(q_sub(q_mul(R, R), J), q_add(V, V));
  const Q SJ = q_mul(S1, J);
  o.y = q_sub(q_mul(R, q_sub(V, o.x)), q_add(SJ, SJ));
  const Q zz = q_add(p.z, q.z);
  o.z = q_mul(q_sub(q_sub(q_mul(zz, zz), Z1Z1), Z2Z2), H);
  return o;
}

}  // namespace

// Window size and workspace of an n-point MSM.
struct MsmPlan {
  uint32_t n = 0, c = 0, windows = 0, half = 0, run = 0, runs = 0, lo_bits = 0, sort_parts = 0;
  uint32_t seg_code = 0;  // bucket segment lengths (ordinary and heavy buckets), packed for the kernels
  uint64_t max_segs = 0;  // bound on their number
  size_t off_blimbs = 0, off_keys = 0, off_tentry = 0, off_tlo = 0, off_list = 0, off_counts = 0, off_offsets = 0,
         off_pcounts = 0, off_poff = 0, off_pcursor = 0, off_slice_off = 0, off_slice_hist = 0, off_cursor = 0,
         off_seg_local = 0,
         off_psegs = 0, off_pseg_off = 0,
         off_seg_sums = 0, off_seg_bucket = 0, off_seg_hist = 0, off_bin_cursor = 0, off_order_seg = 0,
         off_order_bucket = 0, off_buckets = 0, off_parts = 0, off_wout = 0, bytes = 0;
};

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

MsmPlan msm_plan(uint32_t n) {
  MsmPlan p;
  p.n = n;
  uint32_t lg = 0;
  while ((1ull << (lg + 1)) <= n) ++lg;
  // window width by floor(log2 n), measured on gfx950 (sweeps of c = 4..16
  // at 2^6..2^22 random points): up to 2^19 points the latency-bound
  // running sums and window sums cost about the same at every width while
  // the bucket chains shorten as the windows widen, so wide windows win
  // (c = 15 from 2^14: W = 17 windows, the top one full); widths whose top
  // window holds only a few bits of the 254-bit scalars lose
  static const uint8_t kWidth[23] = {4,  4,  4,  4,  4,  4,  5,  5,  8,  8,  9,  10,   // 2^0 .. 2^11
                                     10, 11, 15, 15, 15, 15, 15, 15, 16, 16, 16};  // 2^12 .. 2^22
  p.c = lg < 23 ? kWidth[lg] : 16u;
  p.windows = (255 + p.c - 1) / p.c;  // signed digits: W * c >= 255 bits cover every scalar below 2^254
  p.half = 1u << (p.c - 1);
  // runs of buckets per running-sum thread pair: the kernel is latency-bound
  // (two additions per bucket in a chain), so runs are as short as the
  // window tree allows: about 2^(c-1) / 256 buckets (two parts per run for
  // each of the window kernel's 512 threads), between 4 and 16
  p.run = std::min(p.half, std::max(4u, std::min(16u, p.half / 256)));
  p.runs = p.half / p.run;
  const size_t nb = (size_t)p.windows * p.half;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += up256(bytes);
    return o;
  };
  // the sort's radix split of a bucket: lo_bits (<= 8) within a partition,
  // the rest selects one of sort_parts partitions per window
  p.lo_bits = msm_lo_bits(p.c);
  p.sort_parts = 1u << (p.c - 1 - p.lo_bits);
  const size_t pairs = (size_t)p.windows * n;
  p.off_blimbs = take((size_t)n * msm_base_limb_bytes());
  p.off_keys = take(pairs * 2);
  p.off_tentry = take(pairs * 4);
  p.off_tlo = take(pairs);
  p.off_list = take(pairs * 4);
  const size_t m = (size_t)p.windows * p.sort_parts;  // partitions; the scans write m + 1 words
  // one zeroed region: partition sizes, the most segments of one bucket, the
  // bucket counts, the segment-length histogram
  const size_t bins = msm_seg_bins();
  p.off_pcounts = take((m + 1 + nb + bins) * 4);
  p.off_counts = p.off_pcounts + (m + 1) * 4;
  p.off_seg_hist = p.off_counts + nb * 4;
  p.off_bin_cursor = take((bins + 1) * 4);
  p.off_offsets = take(nb * 4);
  p.off_cursor = take(nb * 4);
  p.off_poff = take((m + 1) * 4);
  p.off_pcursor = take((m + 1) * 4);
  p.off_slice_off = take((m + 1) * 4);
  p.off_slice_hist = take(msm_max_slices(pairs, (uint32_t)m) * ((size_t)1 << p.lo_bits) * 4);
  p.off_seg_local = take(nb * 4);
  p.off_psegs = take(m * 4);
  p.off_pseg_off = take((m + 1) * 4);
  p.seg_code = msm_seg_code(pairs, nb);
  p.max_segs = msm_max_segments(pairs, nb, p.seg_code);
  p.off_seg_sums = take(p.max_segs * msm_point_limb_bytes());
  p.off_seg_bucket = take(p.max_segs * 4);
  p.off_order_seg = take(p.max_segs * 4);
  p.off_order_bucket = take(p.max_segs * 4);
  p.off_buckets = take(nb * msm_point_limb_bytes());
  p.off_parts = take((size_t)p.windows * p.runs * 2 * msm_point_limb_bytes());  // two parts per run
  p.off_wout = take((size_t)p.windows * 96);
  p.bytes = off;
  return p;
}

// Enqueue the device part of an MSM on s; the window sums (W x 96 bytes of
// halo2 words X, Y, Z) land at ws + off_wout.
int msm_enqueue(Engine* eng, const MsmPlan& p, const void* d_scalars, const void* d_bases, void* ws,
                hipStream_t s) {
  char* b = (char*)ws;
  const uint32_t nb = p.windows * p.half;
  uint32_t* counts = (uint32_t*)(b + p.off_counts);  // zeroed with the partition sizes, then counted by the sort
  const bool timing = eng->pass_timing();
  hipError_t e = hipSuccess;
  // one launch (or launch group), bracketed by events while timing is on
  auto step = [&](const char* name, auto&& launch) {
    if (e != hipSuccess) return;
    hipEvent_t ea = nullptr, eb = nullptr;
    if (timing) {
      eng->timed_events(&ea, &eb);
      if (ea) (void)hipEventRecord(ea, s);
    }
    e = launch();
    if (timing) {
      if (eb) (void)hipEventRecord(eb, s);
      eng->finish_timed(ea, eb, name);
    }
  };
  uint32_t* pcounts = (uint32_t*)(b + p.off_pcounts);
  uint32_t* seg_local = (uint32_t*)(b + p.off_seg_local);
  uint32_t* pseg_off = (uint32_t*)(b + p.off_pseg_off);
  const uint32_t m = p.windows * p.sort_parts;
  uint32_t* seg_max = pcounts + m;
  uint32_t* seg_bucket = (uint32_t*)(b + p.off_seg_bucket);
  step("reset_counters_kernel", [&] { return launch_reset_counters(pcounts, m + 1 + nb + msm_seg_bins(), s); });
  step("msm_digits_kernel", [&] {
    return launch_msm_digits(d_scalars, d_bases, p.n, p.c, p.windows, p.lo_bits, b + p.off_blimbs,
                             (uint16_t*)(b + p.off_keys), pcounts, s);
  });
  step("msm_sort (6 kernels)", [&] {
    return launch_msm_sort((const uint16_t*)(b + p.off_keys), p.n, p.c, p.windows, p.lo_bits, pcounts,
                           (uint32_t*)(b + p.off_poff), (uint32_t*)(b + p.off_pcursor),
                           (uint32_t*)(b + p.off_slice_off), (uint32_t*)(b + p.off_tentry), (uint8_t*)(b + p.off_tlo),
                           counts, (uint32_t*)(b + p.off_offsets), (uint32_t*)(b + p.off_cursor),
                           (uint32_t*)(b + p.off_list), (uint32_t*)(b + p.off_slice_hist), p.seg_code, seg_local,
                           (uint32_t*)(b + p.off_psegs), pseg_off, seg_max, s);
  });
  uint32_t* order_seg = (uint32_t*)(b + p.off_order_seg);
  uint32_t* order_bucket = (uint32_t*)(b + p.off_order_bucket);
  step("msm_seg_order (3 kernels)", [&] {
    return launch_msm_seg_order(pseg_off, seg_local, counts, m, p.lo_bits, p.seg_code, nb, p.max_segs,
                                (uint32_t*)(b + p.off_seg_hist), (uint32_t*)(b + p.off_bin_cursor), order_seg,
                                order_bucket, s);
  });
  step("msm_buckets_kernel", [&] {
    return launch_msm_buckets(b + p.off_blimbs, (const uint32_t*)(b + p.off_offsets), counts,
                              (const uint32_t*)(b + p.off_list), pseg_off, seg_local, m, p.lo_bits, p.seg_code, nb,
                              p.max_segs, order_seg, order_bucket, b + p.off_seg_sums, seg_bucket,
                              b + p.off_buckets, s);
  });
  step("msm_bucket_merge (tree levels, gather)", [&] {
    return launch_msm_bucket_merge(b + p.off_seg_sums, seg_bucket, counts, pseg_off, seg_local, seg_max, m,
                                   p.lo_bits, p.seg_code, p.n, nb, p.max_segs, b + p.off_buckets, s);
  });
  step("msm_chunks_kernel",
       [&] { return launch_msm_chunks(b + p.off_buckets, p.half, p.run, p.runs, p.windows, b + p.off_parts, s); });
  step("msm_windows_kernel",
       [&] { return launch_msm_windows(b + p.off_parts, p.runs, p.windows, b + p.off_wout, s); });
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return ZK_ERR_HIP;
  }
  return ZK_OK;
}

// Horner over the window sums (halo2 words, 12 u64 each), added to acc.
void msm_combine(const MsmPlan& p, const uint64_t* wsum, Jac& total) {
  Jac acc = j_identity();
  for (uint32_t w = p.windows; w-- > 0;) {
    for (uint32_t k = 0; k < p.c; ++k) acc = j_double(acc);
    Jac s;
    memcpy(s.x.w, wsum + 12 * w, 32);
    memcpy(s.y.w, wsum + 12 * w + 4, 32);
    memcpy(s.z.w, wsum + 12 * w + 8, 32);
    acc = j_add(acc, s);
  }
  total = j_add(total, acc);
}

// out: X, Y, Z with Z = 1, or all zero for the identity.
void msm_affine(const Jac& acc, uint64_t out[12]) {
  memset(out, 0, 96);
  if (j_is_identity(acc)) return;
  const Q zi = q_inv(acc.z), zi2 = q_mul(zi, zi), zi3 = q_mul(zi2, zi);
  const Q x = q_mul(acc.x, zi2), y = q_mul(acc.y, zi3);
  memcpy(out, x.w, 32);
  memcpy(out + 4, y.w, 32);
  memcpy(out + 8, QONE.w, 32);
}

// An MSM of more than 2^27 points runs as pieces of at most that many, one
// device pass each, their sums added on the host: a pass then has at most
// 16 x 2^27 = 2^31 (window, point) pairs, so every 32-bit position, total and
// loop bound of the sort stays far from wrapping.
constexpr uint64_t MSM_MAX_PIECE = 1ull << 27;
// The host-buffer entry point stages and sums pieces of at most 2^24 points,
// so that its staging buffer and workspace (about 6 GB) stay within the
// host-call HBM budget at any length.
constexpr uint64_t MSM_HOST_PIECE = 1ull << 24;
struct MsmPieces {
  uint64_t count = 0, size = 0;  // pieces of `size` points, the last one shorter
  size_t bytes = 0;              // workspace for the largest of them
};
MsmPieces msm_pieces(uint64_t length, uint64_t max_piece = MSM_MAX_PIECE) {
  MsmPieces q;
  q.count = (length + max_piece - 1) / max_piece;
  q.size = (length + q.count - 1) / q.count;
  const uint64_t last = length - (q.count - 1) * q.size;
  q.bytes = std::max(msm_plan((uint32_t)q.size).bytes, msm_plan((uint32_t)last).bytes);
  return q;
}

// The pieces of an MSM on device operands: each enqueued on s, its window
// sums downloaded and combined into total.
int msm_run(Engine* e, const char* d_scalars, const char* d_bases, uint64_t length, void* ws, hipStream_t s,
            Jac& total) {
  const MsmPieces q = msm_pieces(length);
  total = j_identity();
  for (uint64_t off = 0; off < length; off += q.size) {
    const MsmPlan p = msm_plan((uint32_t)std::min(q.size, length - off));
    int st = msm_enqueue(e, p, d_scalars + off * 32, d_bases + off * 64, ws, s);
    std::vector<uint64_t> wsum((size_t)p.windows * 12);
    if (st == ZK_OK &&
        hipMemcpyAsync(wsum.data(), (char*)ws + p.off_wout, wsum.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
      st = ZK_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && st == ZK_OK) st = ZK_ERR_HIP;
    if (st) return st;
    msm_combine(p, wsum.data(), total);
  }
  return ZK_OK;
}

}  // namespace zkgpu

using namespace zkgpu;

extern "C" {

int zkgpu_msm_device(int device_id, const void* d_scalars, const void* d_bases, uint64_t length, uint64_t* acc,
                     void* stream) {
  if (!acc || length > (1ull << 28)) return ZK_ERR_INVALID;
  if (length && (!d_scalars || !d_bases)) return ZK_ERR_INVALID;
  memset(acc, 0, 96);
  if (length == 0) return ZK_OK;
  int st = ZK_OK;
  Engine* e = Engine::get(device_id, &st);
  if (!e) return st;
  hipStream_t s = (hipStream_t)stream;
  WsLease lease = e->lease_workspace(s, msm_pieces(length).bytes, WS_MSM, &st);
  if (st) return st;
  Jac total;
  st = msm_run(e, (const char*)d_scalars, (const char*)d_bases, length, lease.ptr(), s, total);
  if (st) return st;
  msm_affine(total, acc);
  return ZK_OK;
}

uint64_t parallel_msm_api(size_t device_id, const uint64_t* data, const void* bases, uint64_t length, void* acc) {
  if (!acc || length > (1ull << 28)) return ZK_ERR_INVALID;
  if (length && (!data || !bases)) return ZK_ERR_INVALID;
  if (length == 0) {
    memset(acc, 0, 96);
    return ZK_OK;
  }
  int st = ZK_OK;
  Engine* e = Engine::get((int)device_id, &st);
  if (!e) return (uint64_t)st;
  const MsmPieces q = msm_pieces(length, MSM_HOST_PIECE);
  HostLease ctx(e, q.size * 96, 0, q.bytes);
  if (ctx.status()) return (uint64_t)ctx.status();
  hipStream_t s = ctx->stream;
  char* io = (char*)ctx->io->ptr;
  Jac total = j_identity();
  for (uint64_t off = 0; off < length; off += q.size) {
    const uint64_t m = std::min(q.size, length - off);
    st = stage_h2d((int)device_id, s, io, data + off * 4, m * 32);
    if (st == ZK_OK) st = stage_h2d((int)device_id, s, io + q.size * 32, (const char*)bases + off * 64, m * 64);
    Jac part;
    if (st == ZK_OK) st = msm_run(e, io, io + q.size * 32, m, ctx->aux->ptr, s, part);
    if (st) {
      (void)hipStreamSynchronize(s);
      return (uint64_t)st;
    }
    total = j_add(total, part);
  }
  msm_affine(total, (uint64_t*)acc);
  return ZK_OK;
}

int zkgpu_fill_random_points(int device_id, void* d_out, uint64_t count, uint64_t seed, uint64_t start,
                             void* stream) {
  if (!d_out && count) return ZK_ERR_INVALID;
  int st = ZK_OK;
  if (!Engine::get(device_id, &st)) return st;
  if (count == 0) return ZK_OK;
  return launch_fill_random_points(d_out, count, seed, start, (hipStream_t)stream) == hipSuccess ? ZK_OK : ZK_ERR_HIP;
}

uint64_t zkgpu_msm_workspace_bytes(uint64_t length, uint32_t* window_bits) {
  if (length == 0 || length > (1ull << 28)) return 0;
  const MsmPieces q = msm_pieces(length);
  if (window_bits) *window_bits = msm_plan((uint32_t)q.size).c;
  return q.bytes;
}

}  // extern "C"
