/*
 * zk_gpu_amd — BN256 Fr number-theoretic transform on AMD MI355X (gfx950).
 *
 * C-ABI of libzkgpu_ntt.so. Plain pointers and sizes only.
 *
 * Field elements: BN256 scalar field Fr, r = 0x30644e72e131a029b85045b68181585d
 * 2833e84879b9709143e1f593f0000001, in Montgomery form x * 2^256 mod r, as four
 * little-endian uint64 words per element (halo2's Fr layout; the reference's
 * BigNum, cuda/bn256.h:10-13). Inputs must be canonical (< r); outputs are
 * canonical and bit-identical to halo2's best_fft on the same inputs.
 *
 * Return codes: 0 on success, otherwise one of ZKGPU_ERR_*. No entry point
 * prints or exits (the reference prints and calls exit() on any CUDA error,
 * cuda/fft.cu:379-395).
 */
#ifndef ZKGPU_NTT_H
#define ZKGPU_NTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZKGPU_OK 0
#define ZKGPU_ERR_INVALID 1      /* bad argument (size, pointer, op, device) */
#define ZKGPU_ERR_UNSUPPORTED 2  /* reserved: no entry point returns it (kept so codes stay stable) */
#define ZKGPU_ERR_HIP 3          /* HIP runtime error */
#define ZKGPU_ERR_NOMEM 4        /* device allocation failed */
#define ZKGPU_ERR_NONCANONICAL 5 /* an input element is >= r (host-buffer entry points, validation on) */
#define ZKGPU_ERR_BUDGET 6       /* a job exceeds the serving HBM budget */
#define ZKGPU_ERR_SHUTDOWN 7     /* the server is stopping */
#define ZKGPU_ERR_SCRATCH 8      /* size beyond the no-allocation entry point's scratch limit */

#define ZKGPU_OP_FWD 0       /* X_k = sum_i a_i w^(ik) */
#define ZKGPU_OP_INV 1       /* X_k = divisor * sum_i a_i w^(ik)  (w = omega^-1 for halo2's ifft) */
#define ZKGPU_OP_COSET_FWD 2 /* X_k = sum_i a_i c^i w^(ik) */
#define ZKGPU_OP_COSET_INV 3 /* X_k = divisor * c^k * sum_i a_i w^(ik)  (the inverse of COSET_FWD with c -> c^-1) */

/* ---------------------------------------------------------------------
 * Drop-in entry points of the reference FFI (src/api.rs:3-13,
 * implemented there by cuda/fft.cu:553-584). Same names and argument
 * meaning; widths follow the Rust declarations (usize / u64), which is what
 * the Rust caller passes. Host buffers, synchronous, data transformed in
 * place. `twiddle`/`twiddle_size` (the caller-built lookup table,
 * src/main.rs:128-138) are accepted and ignored: the library keeps its own
 * tables resident on the device, keyed by omega. n must equal 2^log_n,
 * 1 <= log_n <= 28 (the reference requires log_n >= 10).
 * --------------------------------------------------------------------- */

/* replaces do_fft_ffi (src/api.rs:4, cuda/fft.cu:554-558): device smoke
 * check; runs a 2^10 transform round trip on the device, 0 when it holds. */
uint64_t do_fft_ffi(size_t device_id);

/* replaces parallel_fft_api (src/api.rs:6, cuda/fft.cu:560-564):
 * req <- FFT_omega(req) (halo2 best_fft). */
uint64_t parallel_fft_api(size_t device_id, uint64_t* req, const uint64_t* omega, const uint64_t* twiddle,
                          uint64_t n, uint64_t log_n, uint64_t twiddle_size);

/* replaces parallel_ifft_api (src/api.rs:8, cuda/fft.cu:566-570):
 * req <- divisor * FFT_omega(req); the caller passes omega^-1 and n^-1. */
uint64_t parallel_ifft_api(size_t device_id, uint64_t* req, const uint64_t* omega, const uint64_t* twiddle,
                           uint64_t n, uint64_t log_n, uint64_t twiddle_size, const uint64_t* divisor);

/* replaces parallel_coeff_to_extended_part_api (src/api.rs:10,
 * cuda/fft.cu:572-576): req[i] *= c^i, then req <- FFT_omega(req). */
uint64_t parallel_coeff_to_extended_part_api(size_t device_id, uint64_t* req, const uint64_t* omega,
                                             const uint64_t* twiddle, uint64_t n, uint64_t log_n,
                                             uint64_t twiddle_size, const uint64_t* c);

/* replaces parallel_msm_api (src/api.rs:12, cuda/fft.cu:450-551,578-582):
 * *acc <- sum_i data[i] * bases[i] over BN256 G1 (y^2 = x^3 + 3). data: length
 * Fr scalars (4 u64 words each, halo2 Montgomery form, used as the canonical
 * integers they stand for); bases: length affine points (GpuAffine: x then y,
 * 4 u64 words each, Fq in Montgomery form x * 2^256 mod q; x = y = 0 is the
 * identity, as in halo2); acc: one GpuProjective (x, y, z, 12 u64 words,
 * Jacobian: affine = (x / z^2, y / z^3)), written with z = 1 (Montgomery one)
 * or as all zero for the identity. Host buffers, synchronous. The reference's
 * C definition takes a sixth argument that its Rust declaration does not pass;
 * this one follows the Rust declaration. length <= 2^28. */
uint64_t parallel_msm_api(size_t device_id, const uint64_t* data, const void* bases, uint64_t length, void* acc);

/* ---------------------------------------------------------------------
 * Extensions (no counterpart in the reference FFI).
 * --------------------------------------------------------------------- */

/* MSM on device-resident scalars (length x 32 bytes) and bases (length x 64
 * bytes), as parallel_msm_api; acc is a host pointer to 12 words, written
 * before the call returns (the call synchronises `stream`). */
int zkgpu_msm_device(int device_id, const void* d_scalars, const void* d_bases, uint64_t length, uint64_t* acc,
                     void* stream);
/* Seeded synthetic bases: point i of d_out (64 bytes, affine x then y, Fq
 * Montgomery words) = k G for the generator G = (1, 2), k a nonzero 64-bit
 * integer derived from (seed, start + i) by SplitMix64. Asynchronous. */
int zkgpu_fill_random_points(int device_id, void* d_out, uint64_t count, uint64_t seed, uint64_t start,
                             void* stream);
/* Device workspace bytes of a length-point MSM; *window_bits gets the
 * Pippenger window width c (signed digits, 2^(c-1) buckets per window). */
uint64_t zkgpu_msm_workspace_bytes(uint64_t length, uint32_t* window_bits);

/* Coset inverse on host buffers: req <- divisor * c^k * FFT_omega(req)[k];
 * with omega^-1, n^-1, g^-1 it inverts parallel_coeff_to_extended_part_api
 * (coset g). Same size rules as the drop-in entry points. */
uint64_t zkgpu_coset_ifft_api(size_t device_id, uint64_t* req, const uint64_t* omega, uint64_t n,
                              uint64_t log_n, const uint64_t* divisor, const uint64_t* c);

/* Device-resident transform, asynchronous on `stream` (a hipStream_t, NULL
 * for the null stream). d_in and d_out are device pointers to 2^log_n
 * elements; they may be equal (in place). omega is required; divisor is
 * required for OP_INV and OP_COSET_INV; c for the coset ops (host pointers to
 * one element each, read before the call returns). */
int zkgpu_ntt_device(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op,
                     const uint64_t* omega, const uint64_t* divisor, const uint64_t* c, void* stream);

/* zkgpu_ntt_device without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (work under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and results as zkgpu_ntt_device, but the inter-pass
 * workspace and the twiddle/coset tables live in a fixed scratch array of
 * the library (zkgpu_noalloc_scratch_bytes, 120 MiB), held by each GPU the
 * library runs on until the process ends; no hipMalloc, hipFree or other
 * allocating call is made. log_n <= 20 (*max_log_n of
 * zkgpu_noalloc_scratch_bytes); larger sizes return ZKGPU_ERR_SCRATCH. The
 * scratch keeps the last call's tables, reused by the next call with the same
 * size, op and constants. All calls on one device share its scratch: they
 * must be ordered (one stream, or synchronised); calls that overlap race. */
int zkgpu_ntt_device_noalloc(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op,
                             const uint64_t* omega, const uint64_t* divisor, const uint64_t* c, void* stream);
/* Bytes of that scratch array per device; *max_log_n (optional) gets the
 * largest log_n the no-allocation entry point accepts. */
uint64_t zkgpu_noalloc_scratch_bytes(uint32_t* max_log_n);

/* Device-resident transform with halo2's standard domain of size 2^log_n:
 * omega = 7^((r-1)/2^28) squared down to order 2^log_n, divisor = n^-1 and
 * coset generator g = 7 (OP_COSET_INV uses g^-1, so it inverts OP_COSET_FWD). */
int zkgpu_ntt_std_device(int device_id, const void* d_in, void* d_out, uint32_t log_n, int op, void* stream);

/* Domain constants of the standard domain (Montgomery words): which = 0 omega,
 * 1 omega^-1, 2 n^-1, 3 g, 4 g^-1. */
int zkgpu_domain_constant(uint32_t log_n, int which, uint64_t* out4);

/* Build (or fetch) every resident table a transform needs, so the first
 * timed call does not pay for it. Arguments as zkgpu_ntt_device. */
int zkgpu_prepare(int device_id, uint32_t log_n, int op, const uint64_t* omega, const uint64_t* divisor,
                  const uint64_t* c);

/* Seeded synthetic canonical elements: element i of d_out gets synthetic
 * element (start + i) of `seed` (SplitMix64, rejection below r, Montgomery
 * encoded). Asynchronous on `stream`. */
int zkgpu_fill_random(int device_id, void* d_out, uint64_t count, uint64_t seed, uint64_t start, void* stream);

/* Set *d_flag (device uint32) to nonzero when some element is >= r. */
int zkgpu_check_canonical(int device_id, const void* d_in, uint64_t count, uint32_t* d_flag, void* stream);

/* Pass split used for 2^log_n: fills digits[3] (log2 DFT length per pass)
 * and log_cols[3]; returns the number of passes, or 0 when log_n is out of
 * range. Host only. */
int zkgpu_plan_get(uint32_t log_n, uint32_t* digits, uint32_t* log_cols);
/* Install a pass split for 2^log_n (the autotuner's choice); returns
 * ZKGPU_ERR_INVALID for an inconsistent split. npass = 0 drops the installed split (back to the tuned plan file, else the built-in default). */
int zkgpu_plan_set(uint32_t log_n, uint32_t npass, const uint32_t* digits, const uint32_t* log_cols);

/* Same for the two plan tables: kind 0 = single transforms (tiles up to 2^12
 * elements), kind 1 = batched transforms (tiles up to 2^11 elements). */
int zkgpu_plan_get_kind(uint32_t log_n, int kind, uint32_t* digits, uint32_t* log_cols);
int zkgpu_plan_set_kind(uint32_t log_n, int kind, uint32_t npass, const uint32_t* digits, const uint32_t* log_cols);

/* ---------------------------------------------------------------------
 * Batched transforms (no counterpart in the reference): many device-resident
 * polynomials of uneven sizes and mixed ops run as one persistent-kernel
 * launch per pass level, the levels captured once in a hipGraph and replayed
 * by every zkgpu_batch_launch. Each job's result is bit-identical to the
 * same job run alone.
 * --------------------------------------------------------------------- */

#define ZKGPU_JOB_STD_DOMAIN 1u   /* ignore omega/divisor/coset: standard domain of size 2^log_n, g = 7 */
#define ZKGPU_JOB_HOST_BUFFERS 2u /* server jobs only: d_in / d_out are host buffers (see zkgpu_server_submit) */

typedef struct zkgpu_job {
  const void* d_in;     /* device pointer, 2^log_n elements */
  void* d_out;          /* device pointer, may equal d_in */
  uint32_t log_n;
  int32_t op;           /* ZKGPU_OP_* */
  uint32_t flags;       /* ZKGPU_JOB_* */
  uint32_t reserved;
  uint64_t omega[4];    /* as zkgpu_ntt_device */
  uint64_t divisor[4];
  uint64_t coset[4];
} zkgpu_job;

/* Plan a batch, build (or fetch) its resident tables and workspaces and
 * capture its launch graph. The job array is copied; the device buffers it
 * names must stay valid until the batch is destroyed. */
int zkgpu_batch_create(int device_id, const zkgpu_job* jobs, uint32_t count, uint64_t* handle);
/* Enqueue one run of the batch on `stream` (asynchronous). */
int zkgpu_batch_launch(uint64_t handle, void* stream);
/* Pass levels, total tiles and total elements of a batch. */
int zkgpu_batch_info(uint64_t handle, uint32_t* levels, uint32_t* tiles, uint64_t* elements);
/* Wait for the device and free the batch. */
int zkgpu_batch_destroy(uint64_t handle);
/* create + launch + synchronize + destroy. */
int zkgpu_ntt_batch_device(int device_id, const zkgpu_job* jobs, uint32_t count, void* stream);

/* Extended-domain parts of one polynomial (the caller pattern of the
 * reference's parallel_coeff_to_extended_part_api, cuda/fft.cu:572-576,
 * called once per part from src/main.rs:141-148): part j =
 * FFT_omega(a[i] * cosets[j]^i) for j < count, all parts in one batched
 * launch per pass level; the input is read from the device and never copied.
 * cosets holds count elements (4 words each, Montgomery). d_outs[j] must not
 * alias d_in. Runs on `stream` and returns once the parts are written. */
#define ZKGPU_MAX_PARTS 64
int zkgpu_coeff_to_extended_parts_device(int device_id, const void* d_in, void* const* d_outs, uint32_t log_n,
                                         const uint64_t* omega, const uint64_t* cosets, uint32_t count,
                                         void* stream);
/* Same on host buffers: the coefficients are uploaded once, parts (count x n
 * elements, part j at parts + 4 * n * j) are downloaded together. Replaces
 * count calls of parallel_coeff_to_extended_part_api, each of which uploads
 * and downloads the whole polynomial. */
uint64_t zkgpu_coeff_to_extended_parts_api(size_t device_id, const uint64_t* coeffs, uint64_t* parts,
                                           const uint64_t* omega, uint64_t n, uint64_t log_n,
                                           const uint64_t* cosets, uint64_t count);

/* ---------------------------------------------------------------------
 * Serving (replaces the reference's per-call admission, GpuLock in
 * src/gpulock.rs:133-226, for one device): a request loop that takes NTT
 * jobs as they arrive and runs them with continuous batching. One
 * scheduler thread per server issues iterations on its own stream; each
 * iteration is one persistent-kernel launch that runs the next pass of
 * every admitted job, so a job submitted while others are on the device
 * joins at the next iteration and leaves after its last pass. Jobs are
 * admitted in arrival order while their device footprint (output buffer
 * when the server provides it, plus the 36-byte-per-element workspace of a
 * multi-pass plan) fits the HBM budget and fewer than max_active jobs are
 * admitted. Twiddle tables stay resident across jobs. Each result is
 * bit-identical to the same job run alone.
 * --------------------------------------------------------------------- */

typedef struct zkgpu_server_result {
  int32_t status;       /* ZKGPU_OK when done, ZKGPU_ERR_* on failure */
  int32_t done;         /* 1 once the result is in d_out */
  void* d_out;          /* the job's d_out, or the server-owned result buffer */
  uint64_t t_submit_ns; /* steady-clock times: submission, admission, first */
  uint64_t t_admit_ns;  /* launch of one of its passes, completion seen by */
  uint64_t t_start_ns;  /* the scheduler */
  uint64_t t_done_ns;
  uint32_t iterations;    /* server iterations (launches) that ran a pass of the job */
  uint32_t iter_jobs_max; /* most jobs that shared one of those iterations */
} zkgpu_server_result;

typedef struct zkgpu_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t pending, active;     /* waiting for admission / admitted, not done */
  uint64_t iterations;          /* persistent-kernel launches so far */
  uint64_t budget_bytes, in_use_bytes, peak_in_use_bytes, pool_bytes;
  uint64_t elements_done;
} zkgpu_server_stats;

/* Start a server on a device. budget_bytes = 0 selects 64 GiB; max_active =
 * 0 selects 256 admitted jobs. flags is reserved (0). */
int zkgpu_server_create(int device_id, uint64_t budget_bytes, uint32_t max_active, uint32_t flags,
                        uint64_t* handle);
/* Queue a job (non-blocking). job->d_in must hold its data when the job is
 * submitted (the work producing it complete: the server runs on its own
 * stream) and stay valid and unmodified until the job is done; job->d_out =
 * NULL asks for a server-owned result buffer (valid until
 * zkgpu_server_release). With ZKGPU_JOB_HOST_BUFFERS, d_in and d_out are host
 * buffers (pageable memory is fine; they may be equal): the server takes a
 * device buffer for the job at admission, copies the input in on an upload
 * stream, runs the passes and copies the result to d_out on a download
 * stream; the job is done once d_out holds it, and d_in must stay unmodified
 * until then. Uploads, other jobs' passes and downloads overlap. A host job's
 * footprint is its 32-byte-per-element device buffer plus the workspace. A
 * job whose footprint alone exceeds the budget returns ZKGPU_ERR_BUDGET. */
int zkgpu_server_submit(uint64_t handle, const zkgpu_job* job, uint64_t* ticket);
/* State of a job; wait_ms < 0 waits for completion, 0 polls, > 0 waits at
 * most that long. */
int zkgpu_server_poll(uint64_t handle, uint64_t ticket, int32_t wait_ms, zkgpu_server_result* out);
/* Forget a finished job and return its server-owned result buffer to the
 * pool. */
int zkgpu_server_release(uint64_t handle, uint64_t ticket);
int zkgpu_server_stats_get(uint64_t handle, zkgpu_server_stats* out);
/* Finish the admitted jobs, fail the pending ones with ZKGPU_ERR_SHUTDOWN,
 * stop the scheduler and free the pool. */
int zkgpu_server_destroy(uint64_t handle);

/* Resident table bytes on a device; drop them with zkgpu_drop_tables. */
uint64_t zkgpu_table_bytes(int device_id);
int zkgpu_drop_tables(int device_id);

/* Measured throughput of the 32x32->64 multiply-add (v_mad_u64_u32) on the
 * device, in multiply-adds per second: a register-only probe kernel, 8
 * independent chains per thread, 8 waves per SIMD. The Montgomery products
 * of every transform kernel are made of this instruction; this is the peak
 * the VALU roofline is measured against. */
int zkgpu_probe_mad_rate(int device_id, double* mads_per_s);

/* Input validation of the host-buffer entry points (parallel_*_api,
 * zkgpu_coset_ifft_api): when on, the uploaded elements are checked on the
 * device before the transform, and a call whose input holds an element >= r
 * returns ZKGPU_ERR_NONCANONICAL with the caller's buffer left unchanged. The
 * reference does no validation (cuda/fft.cu:275-405). Off by default; on when
 * the environment sets ZKGPU_VALIDATE=1 or after zkgpu_set_validation(1).
 * Returns the previous setting. The device-resident entry points never
 * validate (they are asynchronous). */
int zkgpu_set_validation(int enable);

/* HBM bytes the host-buffer entry points hold on a device: device buffers of
 * the fixed pool of host-call contexts (ZKGPU_HOST_STREAMS contexts, 4 by
 * default, shared by every caller thread) under the ZKGPU_HOST_BUDGET_MB
 * budget (32 GiB by default; a call larger than the whole budget runs alone).
 * *peak_bytes (optional) gets the most held since the budget was last set. */
uint64_t zkgpu_host_held_bytes(int device_id, uint64_t* peak_bytes);
/* Set that budget (bytes; 0 restores the default); idle contexts release
 * their buffers down to it and calls that would pass it wait for running
 * ones, as the reference's GpuLock makes FFT calls wait for GPU memory
 * (src/gpulock.rs:37-44,133-203). Returns the previous budget. */
uint64_t zkgpu_set_host_budget(int device_id, uint64_t bytes);

/* Test hook: after `skip` more successful device allocations by the library,
 * the next `count` fail as if the device were out of memory. */
int zkgpu_debug_fail_allocs(uint32_t skip, uint32_t count);

/* Per-kernel timing, for measurement: while enabled, every pass that a single
 * transform (zkgpu_ntt_device and the entry points built on it) launches
 * carries a start and a stop event tied to its dispatch on the caller's
 * stream (hipExtLaunchKernel), and each MSM kernel launch is bracketed by
 * events on its stream. zkgpu_pass_times waits for
 * the recorded launches and reports, per kernel, the launch count and the
 * summed duration, then forgets them; *count gets the number of kernels
 * (up to max_out entries are written). */
typedef struct zkgpu_kernel_time {
  char name[48];      /* e.g. "ntt_pass_kernel<11>" */
  uint32_t launches;
  uint32_t reserved;
  double total_ms;
} zkgpu_kernel_time;
int zkgpu_pass_timing(int device_id, int enable);
int zkgpu_pass_times(int device_id, zkgpu_kernel_time* out, uint32_t max_out, uint32_t* count);

/* Text of a return code. */
const char* zkgpu_error_string(int code);

/* Library version, major * 10000 + minor * 100 + patch. */
int zkgpu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ZKGPU_NTT_H */
