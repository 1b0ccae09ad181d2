"""Benchmark of the BN256 Fr NTT on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (the metric's k=20 configuration): one step = forward, inverse
(x n^-1) and coset-forward (g = 7) transforms of 2^20 uniform random field
elements, device resident (inputs already in HBM when the timed region
starts). The same line also reports the 2^24 forward transform.

Prints ONE JSON line. `value` is elements transformed per second over the
timed steps. `roofline` describes the dominant kernel (the NTT pass kernel)
against the binding bound, the VALU integer multiply rate: algorithmic work
is (n/2) log2 n Montgomery products per transform (+ n per scaling), each
136 32x32->64 multiply-adds (8-word CIOS). `peak` is the nominal rate (one
v_mad_u64_u32 per lane per clock: 256 CUs x 64 lanes x 2.4 GHz);
`peak_measured` is the rate a register-only probe of that instruction
reaches on this device, measured in the same run. `roofline_hbm` gives the
same kernel against HBM bandwidth (64 bytes per element per pass).
`cpu_baseline` times the CPU restatement of halo2's best_fft (oracle/) on
the host cores over a bounded sample of the same workload.
"""
import argparse
import json
import os
import re
import sys
import time

PEAK_HBM_GBS = 8000.0
# 256 CUs x 64 lanes x 2.4 GHz; v_mad_u64_u32 issues at the full VALU rate on gfx950
PEAK_MAD_TOPS = 256 * 64 * 2.4e9 / 1e12
MADS_PER_MODMUL = 136
# Fq products of one mixed addition in the MSM's bucket sums (XYZZ, 8M + 2S)
FQ_PRODUCTS_PER_MADD = 10


def modmuls(log_n, op):
    n = 1 << log_n
    m = (n // 2) * log_n
    if op != "fwd":
        m += n  # divisor or coset scaling
    if op == "coset_inv":
        m += n
    return m


def profiled_traffic(kernel, log_n):
    """HBM bytes per launch of `kernel` at 2^log_n (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), from the committed
    traffic summaries in profiles/; None when absent."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    pat = re.compile(r"TRAFFIC kernel=(\S+) k=(\d+) bytes_per_launch=(\d+)")
    try:
        names = sorted(os.listdir(d))
    except OSError:
        return None
    for fn in names:
        if not fn.endswith(".txt"):
            continue
        with open(os.path.join(d, fn)) as f:
            for line in f:
                m = pat.match(line.strip())
                if m and m.group(1) == kernel and int(m.group(2)) == log_n:
                    return int(m.group(3))
    return None


def bench_host_api(Z, log_n=20, reps=5):
    """The drop-in host-buffer entry point (parallel_fft_api: host buffer in,
    transformed in place, copies over the host link included)."""
    import numpy as np
    n = 1 << log_n
    req = np.ascontiguousarray(np.random.default_rng(1).integers(0, 1 << 62, size=(n, 4), dtype=np.uint64))
    req[:, 3] &= (1 << 60) - 1  # below r
    omega = Z.domain_constant(log_n, 0)
    Z.check(Z.parallel_fft_api(0, req, omega, None, n, log_n, 0), "parallel_fft_api")
    t0 = time.perf_counter()
    for _ in range(reps):
        Z.check(Z.parallel_fft_api(0, req, omega, None, n, log_n, 0), "parallel_fft_api")
    s = (time.perf_counter() - t0) / reps
    return {"log_n": log_n, "ms": s * 1e3, "elem_per_s": n / s,
            "note": "host buffer in and out (pageable numpy memory), H2D + transform + D2H per call"}


def passes(Z, log_n):
    npass, digits, cols = Z.plan_get(log_n)
    return npass, digits, cols


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=15.0):
    """halo2 best_fft restated in C (oracle/), timed on every host core this
    process may run on (halo2's log_split = floor(log2 threads)), and on one
    thread."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O  # the checker, used here only as the timed CPU path
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    k = 20
    x = O.gen(2001, 1 << k)
    ops = ("fwd", "inv", "coset_fwd")
    O.run_op("fwd", x, k, threads=cores)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        for op in ops:
            O.run_op(op, x, k, threads=cores)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 20:
            break
    value = steps * len(ops) * (1 << k) / el
    # one thread (log_split = 0: halo2's serial path), one step
    t0 = time.perf_counter()
    for op in ops:
        O.run_op(op, x, k, threads=1)
    value_1 = len(ops) * (1 << k) / (time.perf_counter() - t0)
    return {"value": value, "unit": "elem/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
            "sample": f"{steps} steps of fwd+inv+coset_fwd at 2^20 (halo2 best_fft restated in C, "
                      f"{cores} threads, log_split=floor(log2 {cores}))",
            "value_1_thread": value_1, "sample_1_thread": "1 step of fwd+inv+coset_fwd at 2^20, 1 thread"}


def pass_work(log_n, digits, op_names):
    """Algorithmic Montgomery products of each pass of one step, summed over
    the step's transforms: (n/2) * digit for the pass's radix-2 butterflies,
    plus the coset pre-scale (first pass), the inverse's divisor and the
    coset-inverse post-scale (last pass)."""
    n = 1 << log_n
    work = [0] * len(digits)
    for op in op_names:
        for i, d in enumerate(digits):
            work[i] += (n // 2) * d
        if op in ("coset_fwd",):
            work[0] += n
        if op in ("inv", "coset_inv"):
            work[-1] += n
        if op == "coset_inv":
            work[-1] += n
    return work


def kernel_roofline(Z, times, log_n, digits, op_names, p_mul):
    """Per-kernel achieved multiply-add rate from the pass timing (events tied
    to each pass dispatch), and the dominant kernel's figures."""
    names = {}
    work = pass_work(log_n, digits, op_names)
    # which kernel ran pass i: its tile digit (equal digits share a kernel)
    per_kernel = {}
    for name, (launches, total_ms) in times.items():
        d = int(name.split("<")[1].rstrip(">"))
        mm = sum(w for w, dd in zip(work, digits) if dd == d)
        passes_with_d = sum(1 for dd in digits if dd == d)
        per_kernel[name] = {"launches": launches, "avg_us": total_ms / launches * 1e3,
                            "mads_per_launch": mm * MADS_PER_MODMUL / (len(op_names) * passes_with_d)}
    for name, v in per_kernel.items():
        v["achieved_tmad_s"] = v["mads_per_launch"] / (v["avg_us"] * 1e-6) / 1e12
        v["frac"] = v["achieved_tmad_s"] / PEAK_MAD_TOPS
        v["frac_measured"] = v["achieved_tmad_s"] / p_mul
        names[name] = v
    dom = max(names, key=lambda k: names[k]["avg_us"] * names[k]["launches"])
    return dom, names


def bench_batched(torch, Z, dev, stream, count=64, seed=4, reps=10):
    """64 polynomials, sizes log-uniform in 2^12..2^20, ops uniform over
    {fwd, inv, coset_fwd, coset_inv}: the captured batch vs the plain path
    (one device-resident call per job, synchronised after each call, and the
    same calls issued back to back)."""
    import random
    rnd = random.Random(seed)
    sptr = stream.cuda_stream
    jobs_spec = [(rnd.randint(12, 20), rnd.choice([Z.OP_FWD, Z.OP_INV, Z.OP_COSET_FWD, Z.OP_COSET_INV]))
                 for _ in range(count)]
    xs, ys, jobs = [], [], []
    for i, (k, op) in enumerate(jobs_spec):
        x = torch.empty((1 << k) * 4, dtype=torch.int64, device=dev)
        Z.fill_random(x.data_ptr(), 1 << k, 40_000 + i, stream=sptr)
        y = torch.empty_like(x)
        xs.append(x)
        ys.append(y)
        jobs.append(Z.make_job(x.data_ptr(), y.data_ptr(), k, op))
        Z.prepare_std(k, op)
    elements = sum(1 << k for k, _ in jobs_spec)
    b = Z.Batch(jobs)
    for _ in range(3):  # the first replays of a new graph carry one-off costs
        b.launch(sptr)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    runs = []
    for _ in range(5):  # median of five timed runs of `reps` replays
        e0.record(stream)
        for _ in range(reps):
            b.launch(sptr)
        e1.record(stream)
        torch.cuda.synchronize()
        runs.append(e0.elapsed_time(e1) / reps)
    batch_ms = sorted(runs)[2]

    def plain(sync_each):
        t0 = time.perf_counter()
        for (k, op), x, y in zip(jobs_spec, xs, ys):
            Z.ntt_std_device(x.data_ptr(), y.data_ptr(), k, op, stream=sptr)
            if sync_each:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    plain(True)
    plain_sync_ms = min(plain(True) for _ in range(3))
    plain_async_ms = min(plain(False) for _ in range(3))
    info = b.info()
    b.close()
    return {"config": "64 polynomials, log2 sizes uniform in 12..20, ops uniform over fwd/inv/coset_fwd/coset_inv",
            "elements": elements, "levels": info["levels"], "tiles": info["tiles"],
            "batch_ms": batch_ms, "batch_elem_per_s": elements / (batch_ms * 1e-3),
            "plain_sync_ms": plain_sync_ms, "plain_sync_elem_per_s": elements / (plain_sync_ms * 1e-3),
            "plain_async_ms": plain_async_ms, "plain_async_elem_per_s": elements / (plain_async_ms * 1e-3),
            "speedup_vs_plain_sync": plain_sync_ms / batch_ms, "speedup_vs_plain_async": plain_async_ms / batch_ms}


R_MOD = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001


def _std_cosets(Z, log_n, parts):
    """halo2's extended-domain coset shifts c_j = 7 * omega_ext^j (omega_ext of
    order 2^(log_n + log2 parts)), Montgomery words."""
    import numpy as np
    e = parts.bit_length() - 1
    rm = (1 << 256) % R_MOD
    w = Z.domain_constant(log_n + e, 0)
    wv = sum(int(x) << (64 * i) for i, x in enumerate(w)) * pow(rm, -1, R_MOD) % R_MOD
    out = np.zeros((parts, 4), dtype=np.uint64)
    for j in range(parts):
        v = 7 * pow(wv, j, R_MOD) % R_MOD * rm % R_MOD
        out[j] = [(v >> (64 * i)) & ((1 << 64) - 1) for i in range(4)]
    return out


def bench_parts(torch, Z, dev, stream, log_n=20, parts=4, reps=5):
    """Extended-domain parts of one 2^log_n polynomial: one batched call
    against `parts` separate coset-forward transforms (device-resident), and
    the host-buffer entry point against `parts` calls of
    parallel_coeff_to_extended_part_api (each uploads and downloads)."""
    import numpy as np
    n = 1 << log_n
    sptr = stream.cuda_stream
    cs = _std_cosets(Z, log_n, parts)
    w = Z.domain_constant(log_n, 0)
    x = torch.empty(n * 4, dtype=torch.int64, device=dev)
    Z.fill_random(x.data_ptr(), n, 4242, stream=sptr)
    outs = [torch.empty_like(x) for _ in range(parts)]
    torch.cuda.synchronize()

    def batched():
        Z.coeff_to_extended_parts_device(x.data_ptr(), [o.data_ptr() for o in outs], log_n, w, cs, stream=sptr)

    def separate():
        for j, o in enumerate(outs):
            Z.ntt_device(x.data_ptr(), o.data_ptr(), log_n, Z.OP_COSET_FWD, w, None, cs[j], stream=sptr)
        stream.synchronize()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    res = {"log_n": log_n, "parts": parts, "device_batched_ms": timed(batched), "device_separate_ms": timed(separate)}
    h = np.ascontiguousarray(x.cpu().numpy().view(np.uint64).reshape(n, 4))
    hout = np.zeros((parts, n, 4), dtype=np.uint64)
    Z.coeff_to_extended_parts_api(0, h, w, n, log_n, cs, out=hout)
    t0 = time.perf_counter()
    for _ in range(reps):
        Z.coeff_to_extended_parts_api(0, h, w, n, log_n, cs, out=hout)
    res["host_batched_ms"] = (time.perf_counter() - t0) / reps * 1e3
    copies = [h.copy() for _ in range(parts * reps)]
    t0 = time.perf_counter()
    for i, c in enumerate(copies):
        Z.check(Z.parallel_coeff_to_extended_part_api(0, c, w, None, n, log_n, 0, cs[i % parts]))
    res["host_separate_ms"] = (time.perf_counter() - t0) / reps * 1e3
    return res


# The reference's only recorded MSM timings (a log line in a comment,
# cuda/fft.cu:447-448): commit_lagrange at k = 24 on its GPU path, hardware
# unspecified.
REF_MSM_K24_S = {"base_split=128": 1031.868653704, "base_split=256": 951.633471051}


def bench_msm(torch, Z, dev, stream, p_mul, log_ns=(20, 24)):
    """G1 MSM (zkgpu_msm_device) on device-resident synthetic inputs: random
    Fr scalars, bases k_i G. Wall time per call (the call returns the result
    on the host: launches, window sums, host Horner); kernel durations from
    events around each MSM launch. The bucket kernel's algorithmic work is
    n * W mixed additions (Pippenger, W windows of c bits), each 10 Fq products
    (madd-2008-s in XYZZ coordinates: 8M + 2S) of 136 textbook multiply-adds."""
    import ctypes
    sptr = stream.cuda_stream
    out = []
    for log_n in log_ns:
        n = 1 << log_n
        sc = torch.empty(n * 4, dtype=torch.int64, device=dev)
        bs = torch.empty(n * 8, dtype=torch.int64, device=dev)
        Z.fill_random(sc.data_ptr(), n, 7000 + log_n, stream=sptr)
        Z.fill_random_points(bs.data_ptr(), n, 8000 + log_n, stream=sptr)
        torch.cuda.synchronize()
        cbits = ctypes.c_uint32(0)
        ws = Z.lib().zkgpu_msm_workspace_bytes(n, ctypes.byref(cbits))
        c = cbits.value
        windows = (255 + c - 1) // c
        Z.msm_device(sc.data_ptr(), bs.data_ptr(), n, stream=sptr)  # warm-up (workspace, clocks)
        reps = 5 if log_n <= 20 else 3
        t0 = time.perf_counter()
        for _ in range(reps):
            Z.msm_device(sc.data_ptr(), bs.data_ptr(), n, stream=sptr)
        wall = (time.perf_counter() - t0) / reps
        Z.pass_timing(True)
        Z.pass_times()
        Z.msm_device(sc.data_ptr(), bs.data_ptr(), n, stream=sptr)
        kt = Z.pass_times()
        Z.pass_timing(False)
        kernels = {k: v[1] / v[0] * 1e3 for k, v in kt.items()}  # us per launch
        madds = n * windows
        mads = madds * FQ_PRODUCTS_PER_MADD * MADS_PER_MODMUL
        tb = kernels.get("msm_buckets_kernel", 0) * 1e-6
        ach = mads / tb / 1e12 if tb > 0 else None
        out.append({"log_n": log_n, "points": n, "window_bits": c, "windows": windows, "workspace_bytes": ws,
                    "ms": wall * 1e3, "points_per_s": n / wall, "kernels_us": kernels,
                    "roofline": {"bound": "valu", "kernel": "msm_buckets_kernel", "achieved": ach,
                                 "peak": PEAK_MAD_TOPS, "unit": "Tmad/s",
                                 "frac": ach / PEAK_MAD_TOPS if ach else None,
                                 "frac_measured": ach / p_mul if ach else None,
                                 "work": f"{madds} mixed additions x {FQ_PRODUCTS_PER_MADD} Fq products x 136 multiply-adds"}})
        del sc, bs
    res = {"config": "G1 MSM (BN256, y^2 = x^3 + 3), random Fr scalars, synthetic bases k_i G, device resident",
           "sizes": out, "reference_logged_k24_s": REF_MSM_K24_S,
           "reference_note": "the reference's own log line for commit_lagrange at k=24 (cuda/fft.cu:447-448), "
                             "hardware unspecified; not a same-hardware comparison"}
    k24 = [o for o in out if o["log_n"] == 24]
    if k24:
        res["speedup_vs_reference_logged_k24"] = REF_MSM_K24_S["base_split=256"] / (k24[0]["ms"] * 1e-3)
    return res


def cpu_baseline_msm(budget_s=10.0):
    """The oracle's MSM (bit-serial Straus, every host core) on 2^14 points."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O  # the checker, used here only as the timed CPU path
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    n = 1 << 14
    bases = O.g1_arith_bases(n, 12345, 6789)
    scalars = O.gen(9100, n)
    reps, t0 = 0, time.perf_counter()
    while True:
        O.g1_msm(scalars, bases, threads=cores)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 5:
            break
    el = (time.perf_counter() - t0) / reps
    return {"value": n / el, "unit": "points/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{reps} MSMs of 2^14 points (oracle: bit-serial Straus in C, {cores} threads)"}


def _serve_spec(count, seed=5, lo=14, hi=22):
    import random
    rnd = random.Random(seed)
    return [(rnd.randint(lo, hi), rnd.choice([0, 1, 2, 3])) for _ in range(count)]


def _arrivals(count, rate, seed):
    """Poisson process: exponential inter-arrival times (seconds from start)."""
    import random
    rnd = random.Random(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += rnd.expovariate(rate)
        out.append(t)
    return out


def _wait_until(t_ns):
    while True:
        d = t_ns - time.monotonic_ns()
        if d <= 0:
            return
        time.sleep(d / 2e9 if d > 400_000 else 0)


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(q * len(xs)))]


def bench_serving(torch, Z, dev, stream, count=10_000, loads=(0.5, 0.9, 1.5), seed=5):
    """Poisson stream of `count` jobs, k uniform in 14..22, ops uniform over the
    four ops, device-resident operands (every job reads a prefix of one random
    2^22-element input; results go to server-owned buffers, released on
    completion), HBM admission budget 64 GiB. The plain path serves the same
    arrivals one call at a time (ntt_std_device + synchronize, FIFO)."""
    import threading
    sptr = stream.cuda_stream
    spec = _serve_spec(count)
    elements = sum(1 << k for k, _ in spec)
    src = torch.empty((1 << 22) * 4, dtype=torch.int64, device=dev)
    Z.fill_random(src.data_ptr(), 1 << 22, 8080, stream=sptr)
    outs = {k: torch.empty((1 << k) * 4, dtype=torch.int64, device=dev) for k in range(14, 23)}
    for k in range(14, 23):
        for op in range(4):
            Z.prepare_std(k, op)
    torch.cuda.synchronize()

    def plain_run(arr_ns):
        lat = []
        for (k, op), a in zip(spec, arr_ns):
            if a is not None:
                _wait_until(a)
            Z.ntt_std_device(src.data_ptr(), outs[k].data_ptr(), k, op, stream=sptr)
            stream.synchronize()
            if a is not None:
                lat.append(time.monotonic_ns() - a)
        return lat

    srv = Z.Server(budget_bytes=64 << 30)

    def served_run(arr_ns, check_every=0):
        tickets = [None] * len(spec)
        checked = []
        lat = []
        t_first = [None]

        def submitter():
            for i, ((k, op), a) in enumerate(zip(spec, arr_ns)):
                if a is not None:
                    _wait_until(a)
                tickets[i] = srv.submit(Z.make_job(src.data_ptr(), None, k, op))

        th = threading.Thread(target=submitter)
        t0 = time.monotonic_ns()
        th.start()
        last_done = 0
        for i, ((k, op), a) in enumerate(zip(spec, arr_ns)):
            while tickets[i] is None:
                time.sleep(0.0001)
            r = srv.wait(tickets[i])
            if t_first[0] is None:
                t_first[0] = r.t_submit_ns
            last_done = max(last_done, r.t_done_ns)
            lat.append(r.t_done_ns - (a if a is not None else r.t_submit_ns))
            if check_every and i % check_every == 0:
                checked.append((i, k, op, _device_copy(torch, r.d_out, k)))
            srv.release(tickets[i])
        th.join()
        return lat, (last_done - t0) * 1e-9, checked

    # warm-up: every size class through both paths (pool buffers, tables)
    warm = [None] * len(spec)
    plain_run(warm[:200])
    served_run(warm[:200])
    # plain capacity: all jobs back to back, one call + synchronize each
    t0 = time.perf_counter()
    plain_run([None] * len(spec))
    plain_s = time.perf_counter() - t0
    plain_rate = count / plain_s
    # served capacity: all jobs submitted at once
    lat_sat, served_s, checked = served_run([None] * len(spec), check_every=max(1, count // 40))
    # parity: sampled served results equal the plain path's
    mismatches = 0
    for i, k, op, got in checked:
        Z.ntt_std_device(src.data_ptr(), outs[k].data_ptr(), k, op, stream=sptr)
        stream.synchronize()
        mismatches += int(not torch.equal(got, outs[k]))
    res = {"config": f"{count} jobs, k uniform in 14..22, ops uniform over fwd/inv/coset_fwd/coset_inv, "
                     "Poisson arrivals, HBM admission budget 64 GiB, device-resident operands",
           "elements": elements,
           "plain_elem_per_s": elements / plain_s, "plain_jobs_per_s": plain_rate,
           "served_elem_per_s": elements / served_s, "served_jobs_per_s": count / served_s,
           "speedup_vs_plain": plain_s / served_s,
           "saturated_p50_ms": _pct(lat_sat, 0.5) / 1e6, "saturated_p99_ms": _pct(lat_sat, 0.99) / 1e6,
           "sampled_results_checked": len(checked), "sampled_mismatches": mismatches,
           "loads": []}
    for i, rho in enumerate(loads):
        lam = rho * plain_rate
        arr = _arrivals(count, lam, seed=100 + i)
        base = time.monotonic_ns() + 2_000_000
        arr_ns = [base + int(a * 1e9) for a in arr]
        lat_p = plain_run(arr_ns)
        base = time.monotonic_ns() + 2_000_000
        arr_ns = [base + int(a * 1e9) for a in arr]
        lat_s, s_s, _ = served_run(arr_ns)
        res["loads"].append({"offered_jobs_per_s": lam, "load_vs_plain_capacity": rho,
                             "plain_p50_ms": _pct(lat_p, 0.5) / 1e6, "plain_p99_ms": _pct(lat_p, 0.99) / 1e6,
                             "served_p50_ms": _pct(lat_s, 0.5) / 1e6, "served_p99_ms": _pct(lat_s, 0.99) / 1e6,
                             "served_elem_per_s": elements / s_s})
    # per size class: plain per-call rate against the served rate of a stream
    # of that size only (all jobs submitted at once)
    per_k = []
    for k in range(14, 23):
        jobs = [(k, op % 4) for op in range(48)]
        el = len(jobs) << k
        t0 = time.perf_counter()
        for _, op in jobs:
            Z.ntt_std_device(src.data_ptr(), outs[k].data_ptr(), k, op, stream=sptr)
            stream.synchronize()
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        tks = [srv.submit(Z.make_job(src.data_ptr(), None, k, op)) for _, op in jobs]
        for t in tks:
            srv.wait(t)
            srv.release(t)
        t_served = time.perf_counter() - t0
        per_k.append({"log_n": k, "plain_elem_per_s": el / t_plain, "served_elem_per_s": el / t_served,
                      "ratio": t_plain / t_served, "share_of_mix_elements": (1 << k) / sum(1 << q for q in range(14, 23))})
    res["per_k"] = per_k
    st = srv.stats()
    srv.close()
    res["server"] = {"iterations": st["iterations"], "peak_in_use_bytes": st["peak_in_use_bytes"],
                     "pool_bytes": st["pool_bytes"], "failed": st["failed"]}
    return res


def _host_call(Z, buf, k, op):
    """One drop-in host-buffer call (the reference's per-call path: copy in,
    transform, copy out), in place on buf."""
    n = 1 << k
    inverse = op in (1, 3)
    w = Z.domain_constant(k, 1 if inverse else 0)
    if op == 0:
        rc = Z.parallel_fft_api(0, buf, w, None, n, k, 0)
    elif op == 1:
        rc = Z.parallel_ifft_api(0, buf, w, None, n, k, 0, Z.domain_constant(k, 2))
    elif op == 2:
        rc = Z.parallel_coeff_to_extended_part_api(0, buf, w, None, n, k, 0, Z.domain_constant(k, 3))
    else:
        rc = Z.zkgpu_coset_ifft_api(0, buf, w, n, k, Z.domain_constant(k, 2), Z.domain_constant(k, 4))
    Z.check(rc, "host call")


def bench_host_modes(torch, Z, dev, stream, serve_count=1000, slots=24):
    """Host-buffer modes (data in pageable host memory before and after, as
    the reference's FFI takes it): the plain path is one drop-in call per job
    (parallel_*_api: copy in, transform, copy out); the served path submits the
    same jobs to the server as ZKGPU_JOB_HOST_BUFFERS jobs, which overlaps one
    job's upload with others' passes and downloads. Every job transforms its
    host buffer in place (a ring of `slots` buffers, at most `slots` jobs
    outstanding). Two workloads: the 64-polynomial batch mix (k 12..20) and
    `serve_count` jobs of the serving mix (k 14..22)."""
    import random
    import threading

    import numpy as np
    src = torch.empty((1 << 22) * 4, dtype=torch.int64, device=dev)
    Z.fill_random(src.data_ptr(), 1 << 22, 9090, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    host_src = src.cpu().numpy().view(np.uint64).reshape(-1, 4)
    ring = [np.empty(((1 << 22), 4), dtype=np.uint64) for _ in range(slots)]
    for r in ring:
        np.copyto(r, host_src)  # touch every page once
    for k in range(12, 23):
        for op in range(4):
            Z.prepare_std(k, op)

    def run_plain(spec):
        t0 = time.perf_counter()
        for i, (k, op) in enumerate(spec):
            _host_call(Z, ring[i % slots][:1 << k], k, op)
        return time.perf_counter() - t0

    def run_served(spec):
        srv = Z.Server(budget_bytes=64 << 30)
        sem = threading.Semaphore(slots)
        tickets = [None] * len(spec)
        lat = []

        def waiter():
            for i in range(len(spec)):
                while tickets[i] is None:
                    time.sleep(0.00005)
                r = srv.wait(tickets[i])
                lat.append(r.t_done_ns - r.t_submit_ns)
                srv.release(tickets[i])
                sem.release()

        th = threading.Thread(target=waiter)
        t0 = time.perf_counter()
        th.start()
        for i, (k, op) in enumerate(spec):
            sem.acquire()
            buf = ring[i % slots]
            tickets[i] = srv.submit(Z.make_job(buf.ctypes.data, buf.ctypes.data, k, op, host=True))
        th.join()
        el = time.perf_counter() - t0
        srv.close()
        return el, lat

    out = {}
    rnd = random.Random(4)
    batch_spec = [(rnd.randint(12, 20), rnd.choice([0, 1, 2, 3])) for _ in range(64)]
    serve_spec = _serve_spec(serve_count, seed=6)
    for name, spec, reps in (("batch_mix_64", batch_spec, 3), (f"serving_mix_{serve_count}", serve_spec, 1)):
        elements = sum(1 << k for k, _ in spec)
        run_plain(spec[:8])
        run_served(spec[:8])
        plain_s = min(run_plain(spec) for _ in range(reps))
        served = [run_served(spec) for _ in range(reps)]
        served_s, lat = min(served, key=lambda x: x[0])
        out[name] = {"jobs": len(spec), "elements": elements, "plain_s": plain_s,
                     "plain_elem_per_s": elements / plain_s, "served_s": served_s,
                     "served_elem_per_s": elements / served_s, "speedup_vs_plain": plain_s / served_s,
                     "served_p50_ms": _pct(lat, 0.5) / 1e6, "served_p99_ms": _pct(lat, 0.99) / 1e6,
                     "outstanding_jobs": slots}
    out["note"] = ("host-buffer mode: inputs and results in pageable host memory; plain = one parallel_*_api call "
                   "per job (copy in, transform, copy out); served = ZKGPU_JOB_HOST_BUFFERS jobs, in place")
    return out


def _device_copy(torch, d_ptr, k):
    """Tensor copy of a server-owned result buffer."""
    import ctypes
    n = 1 << k
    t = torch.empty(n * 4, dtype=torch.int64, device="cuda")
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(d_ptr), ctypes.c_size_t(n * 32), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-k24", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer entry point timing")
    ap.add_argument("--no-parts", action="store_true", help="skip the extended-domain parts timing")
    ap.add_argument("--no-msm", action="store_true", help="skip the G1 MSM timing")
    ap.add_argument("--serve-jobs", type=int, default=10_000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 256 MiB once on the device first: a kernel of known traffic for "
                         "calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); "
              "the transform's multi-GPU split is design-only", file=sys.stderr)
        sys.exit(2)

    import torch
    import zk_gpu_amd as Z

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream
    p_mul = max(Z.probe_mad_rate() for _ in range(3)) / 1e12  # measured v_mad_u64_u32 rate, T/s
    if args.copy_probe:
        src = torch.ones(1 << 25, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        del src, dst
    k = 20
    n = 1 << k
    x = torch.empty(n * 4, dtype=torch.int64, device=dev)
    outs = [torch.empty(n * 4, dtype=torch.int64, device=dev) for _ in range(3)]
    Z.fill_random(x.data_ptr(), n, 2001, stream=sptr)
    ops = [("fwd", Z.OP_FWD), ("inv", Z.OP_INV), ("coset_fwd", Z.OP_COSET_FWD)]
    for _, op in ops:
        Z.prepare_std(k, op)

    def step():
        for (_, op), o in zip(ops, outs):
            Z.ntt_std_device(x.data_ptr(), o.data_ptr(), k, op, stream=sptr)

    # Clock ramp: the GPU's clocks settle only after tens of milliseconds of
    # load (20 steps after 5 warmup steps time 0.37 ms per step, 200 after 50
    # time 0.32 ms), so the step runs untimed for 100 ms before the warmup.
    t_ramp = time.perf_counter() + 0.1
    while time.perf_counter() < t_ramp:
        step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    value = 3 * n / (ms * 1e-3)

    npass, digits, cols = passes(Z, k)
    launches = 3 * npass
    launch_ms = ms / launches
    mm_step = sum(modmuls(k, name) for name, _ in ops)
    # the same steps again with start/stop events tied to every pass dispatch
    # (hipExtLaunchKernel, no extra packets): per-kernel durations
    Z.pass_timing(True)
    Z.pass_times()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_timed = e0.elapsed_time(e1) / args.steps
    times = Z.pass_times()
    Z.pass_timing(False)
    dom, kern = kernel_roofline(Z, times, k, digits, [o for o, _ in ops], p_mul)
    kd = kern[dom]
    bytes_per_launch = 64 * n
    hbm = bytes_per_launch / (kd["avg_us"] * 1e-6) / 1e9

    res = {
        "metric": "BN256 Fr NTT elem/s at k=20/24; served mixed-size vs plain; p50/p99 lat",
        "value": value,
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded SplitMix64 uniform Fr elements, Montgomery form, device resident)",
        "config": {"workload": "fwd+inv+coset_fwd NTT over BN256 Fr at n=2^20 (one step = 3 transforms)",
                   "log_n": k, "ops": [o for o, _ in ops], "coset_g": 7,
                   "plan": {"digits": digits, "log_cols": cols}},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": kd["achieved_tmad_s"], "peak": PEAK_MAD_TOPS,
                     "unit": "Tmad/s", "frac": kd["frac"], "traffic": profiled_traffic(dom, k),
                     "peak_measured": p_mul, "frac_measured": kd["frac_measured"],
                     "avg_launch_us": kd["avg_us"], "launches_timed": kd["launches"],
                     "work_per_launch": f"{kd['mads_per_launch']:.4g} 32x32->64 multiply-adds "
                                        f"({kd['mads_per_launch'] / MADS_PER_MODMUL:.4g} Montgomery products x 136)",
                     "timing": "start/stop events tied to each pass dispatch on the bench stream over "
                               f"{args.steps} steps ({ms_timed:.4f} ms per step with the events, "
                               f"{ms:.4f} ms without)",
                     "kernels": kern,
                     "step_average": {"achieved": mm_step * MADS_PER_MODMUL / launches / (launch_ms * 1e-3) / 1e12,
                                      "launch_ms": launch_ms}},
        "roofline_hbm": {"bound": "hbm", "kernel": dom, "achieved": hbm, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": hbm / PEAK_HBM_GBS, "bytes_per_launch": bytes_per_launch},
        "wall_s_timed_region": wall,
    }
    res["roofline"]["step_average"]["frac"] = res["roofline"]["step_average"]["achieved"] / PEAK_MAD_TOPS

    if not args.no_k24:
        k24 = 24
        n24 = 1 << k24
        x24 = torch.empty(n24 * 4, dtype=torch.int64, device=dev)
        y24 = torch.empty(n24 * 4, dtype=torch.int64, device=dev)
        Z.fill_random(x24.data_ptr(), n24, 3001, stream=sptr)
        Z.prepare_std(k24, Z.OP_FWD)
        for _ in range(2):
            Z.ntt_std_device(x24.data_ptr(), y24.data_ptr(), k24, Z.OP_FWD, stream=sptr)
        torch.cuda.synchronize()
        reps = max(3, args.steps // 2)
        runs = []
        for _ in range(5):  # median of five timed runs of `reps` transforms
            e0.record(stream)
            for _ in range(reps):
                Z.ntt_std_device(x24.data_ptr(), y24.data_ptr(), k24, Z.OP_FWD, stream=sptr)
            e1.record(stream)
            torch.cuda.synchronize()
            runs.append(e0.elapsed_time(e1) / reps)
        ms24 = sorted(runs)[2]
        np24, d24, c24 = passes(Z, k24)
        Z.pass_timing(True)
        Z.pass_times()
        for _ in range(reps):
            Z.ntt_std_device(x24.data_ptr(), y24.data_ptr(), k24, Z.OP_FWD, stream=sptr)
        torch.cuda.synchronize()
        t24 = Z.pass_times()
        Z.pass_timing(False)
        dom24, kern24 = kernel_roofline(Z, t24, k24, d24, ["fwd"], p_mul)
        kd24 = kern24[dom24]
        res["k24"] = {"op": "fwd", "ms": ms24, "elem_per_s": n24 / (ms24 * 1e-3),
                      "plan": {"digits": d24, "log_cols": c24},
                      "roofline": {"bound": "valu", "kernel": dom24, "achieved": kd24["achieved_tmad_s"],
                                   "peak": PEAK_MAD_TOPS, "unit": "Tmad/s", "frac": kd24["frac"],
                                   "traffic": profiled_traffic(dom24, k24),
                                   "peak_measured": p_mul, "frac_measured": kd24["frac_measured"],
                                   "avg_launch_us": kd24["avg_us"], "kernels": kern24,
                                   "whole_transform_frac": modmuls(k24, "fwd") * MADS_PER_MODMUL /
                                                           (ms24 * 1e-3) / 1e12 / PEAK_MAD_TOPS},
                      "roofline_hbm": {"bound": "hbm", "achieved": 64 * n24 / (kd24["avg_us"] * 1e-6) / 1e9,
                                       "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                       "frac": 64 * n24 / (kd24["avg_us"] * 1e-6) / 1e9 / PEAK_HBM_GBS}}
        del x24, y24

    if not args.no_batch:
        res["batched"] = bench_batched(torch, Z, dev, stream)

    if not args.no_host:
        res["host_api_k20"] = bench_host_api(Z)
        res["host_buffer_modes"] = bench_host_modes(torch, Z, dev, stream)
    if not args.no_parts:
        res["extended_parts"] = bench_parts(torch, Z, dev, stream)

    if not args.no_serve:
        res["served"] = bench_serving(torch, Z, dev, stream, count=args.serve_jobs)

    if not args.no_msm:
        res["msm"] = bench_msm(torch, Z, dev, stream, p_mul)

    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline()
        if not args.no_msm:
            res["msm"]["cpu_baseline"] = cpu_baseline_msm()
            m20 = [o for o in res["msm"]["sizes"] if o["log_n"] == 20]
            if m20:
                res["msm"]["speedup_vs_cpu_2^20"] = m20[0]["points_per_s"] / res["msm"]["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
