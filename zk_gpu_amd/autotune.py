"""Autotuner for the pass split of the NTT (radix/tile/pass schedule).

For each transform size 2^k it times every candidate split -- number of
passes, the log2 DFT length of each pass (the tile radix) and the log2
columns per tile (the tile width) -- on the device and keeps the fastest.
Results are written as plain text lines

    log_n npass d0 d1 d2 c0 c1 c2 ms

to zk_gpu_amd/plans_gfx950.txt (single transforms, tiles up to 2^12
elements) or zk_gpu_amd/plans_batch_gfx950.txt (batched transforms, tiles up
to 2^11 elements). Batched plans are tuned on a mixed workload (--mix): a
batch of jobs of several sizes and ops is timed as a whole while the split of
one size class at a time is varied (coordinate descent over the classes), which the library reads at start-up (ZKGPU_PLAN_FILE /
ZKGPU_BATCH_PLAN_FILE override the paths).

    python -m zk_gpu_amd.autotune --k 12-24 [--reps 5] [--out PATH]
    python -m zk_gpu_amd.autotune --kind batch --mix mixed64 [--sweeps 2]
"""
import argparse
import itertools
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = {"single": os.path.join(HERE, "plans_gfx950.txt"),
               "batch": os.path.join(HERE, "plans_batch_gfx950.txt")}
TILE_LOG = {"single": 12, "batch": 11}


def candidates(k, min_tiles_log=8, max_passes=3, tile_log_max=12, wide=False):
    """Valid pass splits of 2^k: digits each <= 12, tile N*C <= 4096,
    column limits as in the engine (pass p < P: columns <= 2^(remaining),
    last pass: columns <= 2^digit0). By default every pass aims at one
    common tile size; wide=True lets each pass choose its own."""
    out = []
    TILE_LOG_MAX = tile_log_max
    for npass in range(1, max_passes + 1):
        if npass == 1 and k > TILE_LOG_MAX:
            continue
        for digits in itertools.product(range(1, TILE_LOG_MAX + 1), repeat=npass):
            if sum(digits) != k:
                continue
            # balanced splits only: the modmul count barely depends on the
            # split, while lopsided splits lose tile parallelism
            if npass == 2 and abs(digits[0] - digits[1]) > 4:
                continue
            if npass == 3 and max(digits) - min(digits) > 3:
                continue
            # tile width: one common target tile size for every pass
            # (2^t elements per tile, t in 8..12) or no columns at all
            avails = []
            log_m = k
            for i, d in enumerate(digits):
                if npass == 1:
                    avail = 0
                elif i + 1 < npass:
                    avail = log_m - d
                else:
                    avail = digits[0]
                avails.append(min(avail, TILE_LOG_MAX - d))
                log_m -= d
            seen = set()
            targets = [t for t in (0, 8, 9, 10, 11, 12) if t <= TILE_LOG_MAX]
            if wide:
                per_pass = [sorted(set(min(max(t - d, 0), a) for t in targets)) for d, a in zip(digits, avails)]
                col_sets = itertools.product(*per_pass)
            else:
                col_sets = (tuple(min(max(t - d, 0), a) for d, a in zip(digits, avails)) for t in targets)
            for cols in col_sets:
                cols = tuple(cols)
                if cols in seen:
                    continue
                seen.add(cols)
                # enough tiles per pass to fill the chip (unless the size is small)
                if all((k - d - c) >= min(min_tiles_log, k - d) for d, c in zip(digits, cols)):
                    out.append((list(digits), list(cols)))
    return out


# ops timed per repetition: forward, inverse and coset forward, the mix the
# benchmark's step runs (their first and last passes differ)
TUNE_OPS = (0, 1, 2)


def time_plan(Z, torch, k, digits, cols, x, y, stream, reps, prepared=False):
    """ms per transform (mean over the ops of TUNE_OPS)."""
    if not prepared:
        Z.plan_set(k, digits, cols)
        Z.lib().zkgpu_drop_tables(0)
        for op in TUNE_OPS:
            Z.prepare_std(k, op)
    for op in TUNE_OPS:
        Z.ntt_std_device(x.data_ptr(), y.data_ptr(), k, op, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        for op in TUNE_OPS:
            Z.ntt_std_device(x.data_ptr(), y.data_ptr(), k, op, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps / len(TUNE_OPS)


def confirm(Z, torch, k, finalists, x, y, stream, reps, runs=5):
    """Median over `runs` interleaved timings of each finalist (single-transform
    plans): the sweep's single timing of hundreds of candidates picks noise
    between plans a few percent apart."""
    samples = {i: [] for i in range(len(finalists))}
    for _ in range(runs):
        for i, (digits, cols) in enumerate(finalists):
            Z.plan_set(k, digits, cols)
            Z.lib().zkgpu_drop_tables(0)
            for op in TUNE_OPS:
                Z.prepare_std(k, op)
            samples[i].append(time_plan(Z, torch, k, digits, cols, x, y, stream, reps, prepared=True))
    med = [sorted(v)[len(v) // 2] for _, v in sorted(samples.items())]
    return sorted((m, list(f[0]), list(f[1])) for m, f in zip(med, finalists))


def time_plan_batch(Z, torch, k, digits, cols, xs, ys, stream, reps):
    """ms per job of a batch of same-size forward jobs."""
    Z.plan_set(k, digits, cols, kind=Z.PLAN_BATCH)
    Z.lib().zkgpu_drop_tables(0)
    jobs = [Z.make_job(x.data_ptr(), y.data_ptr(), k, Z.OP_FWD) for x, y in zip(xs, ys)]
    b = Z.Batch(jobs)
    b.launch(stream.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        b.launch(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    b.close()
    return e0.elapsed_time(e1) / reps / len(jobs)


def tune(ks, reps=5, out=None, verbose=True, max_candidates=None, kind="single", wide=False):
    import torch

    import zk_gpu_amd as Z
    out = DEFAULT_OUT[kind] if out is None else out
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    best = {}
    for k in ks:
        n = 1 << k
        count = 1 if kind == "single" else max(1, (1 << 22) >> k)
        xs = [torch.empty(n * 4, dtype=torch.int64, device=dev) for _ in range(count)]
        ys = [torch.empty_like(x) for x in xs]
        for i, x in enumerate(xs):
            Z.fill_random(x.data_ptr(), n, 99 + i, stream=stream.cuda_stream)
        cands = candidates(k, tile_log_max=TILE_LOG[kind], wide=wide)
        if max_candidates:
            cands = cands[:max_candidates]
        results = []
        for digits, cols in cands:
            if kind == "single":
                ms = time_plan(Z, torch, k, digits, cols, xs[0], ys[0], stream, reps)
            else:
                ms = time_plan_batch(Z, torch, k, digits, cols, xs, ys, stream, reps)
            results.append((ms, digits, cols))
            if verbose:
                print(f"k={k} digits={digits} cols={cols} {ms:.5f} ms", flush=True)
        results.sort()
        if kind == "single":
            # the five fastest and the plan in use now, re-timed
            Z.plan_reset(k, kind=Z.PLAN_SINGLE)
            npass, d0, c0 = Z.plan_get(k, Z.PLAN_SINGLE)
            finalists = [(r[1], r[2]) for r in results[:5]]
            if (d0, c0) not in finalists:
                finalists.append((d0, c0))
            results = confirm(Z, torch, k, finalists, xs[0], ys[0], stream, reps * 3)
            if verbose:
                for r in results:
                    print(f"confirm k={k} digits={r[1]} cols={r[2]} {r[0]:.5f} ms", flush=True)
        best[k] = results[0]
        Z.plan_reset(k, kind=Z.PLAN_SINGLE if kind == "single" else Z.PLAN_BATCH)
        if verbose:
            print(f"BEST k={k}: {results[0]}", flush=True)
        del xs, ys
        Z.lib().zkgpu_drop_tables(0)
    if out:
        existing = {}
        if os.path.exists(out):
            for line in open(out):
                f = line.split()
                if len(f) >= 9 and not line.startswith("#"):
                    existing[int(f[0])] = line.rstrip("\n")
        for k, (ms, digits, cols) in best.items():
            d = digits + [0] * (3 - len(digits))
            c = cols + [0] * (3 - len(cols))
            existing[k] = f"{k} {len(digits)} {d[0]} {d[1]} {d[2]} {c[0]} {c[1]} {c[2]} {ms:.5f}"
        with open(out, "w") as f:
            what = "forward transform" if kind == "single" else "batched forward transforms, ms per job"
            f.write(f"# log_n npass d0 d1 d2 c0 c1 c2 ms  (autotuned on gfx950, {what})\n")
            for k in sorted(existing):
                f.write(existing[k] + "\n")
    return best


MIXES = {
    # 64 polynomials, log2 sizes uniform in 12..20, ops uniform over the four ops
    "mixed64": dict(count=64, lo=12, hi=20, seed=4),
    # serving-like jobs, log2 sizes uniform in 14..22
    "serve": dict(count=48, lo=14, hi=22, seed=5),
}


def tune_batch_mix(mix="mixed64", reps=3, sweeps=2, out=None, verbose=True, wide=False):
    """Coordinate descent over the size classes of a mixed batch: for each
    class in turn, try every candidate split with the other classes fixed and
    keep the one that makes the whole batch fastest."""
    import random

    import torch

    import zk_gpu_amd as Z
    out = DEFAULT_OUT["batch"] if out is None else out
    cfg = MIXES[mix]
    rnd = random.Random(cfg["seed"])
    spec = [(rnd.randint(cfg["lo"], cfg["hi"]), rnd.choice([0, 1, 2, 3])) for _ in range(cfg["count"])]
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    xs, ys = [], []
    for i, (k, _) in enumerate(spec):
        x = torch.empty((1 << k) * 4, dtype=torch.int64, device=dev)
        Z.fill_random(x.data_ptr(), 1 << k, 70_000 + i, stream=stream.cuda_stream)
        xs.append(x)
        ys.append(torch.empty_like(x))
    jobs = [Z.make_job(x.data_ptr(), y.data_ptr(), k, op) for (k, op), x, y in zip(spec, xs, ys)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def run():
        b = Z.Batch(jobs)
        b.launch(stream.cuda_stream)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            b.launch(stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        b.close()
        return e0.elapsed_time(e1) / reps

    classes = sorted(set(k for k, _ in spec))
    current = {k: Z.plan_get(k, Z.PLAN_BATCH)[1:] for k in classes}
    best_ms = run()
    if verbose:
        print(f"start {best_ms:.4f} ms", flush=True)
    for sweep in range(sweeps):
        for k in classes:
            for digits, cols in candidates(k, tile_log_max=TILE_LOG["batch"], wide=wide):
                Z.plan_set(k, digits, cols, kind=Z.PLAN_BATCH)
                ms = run()
                if ms < best_ms:
                    best_ms = ms
                    current[k] = (digits, cols)
                    if verbose:
                        print(f"sweep {sweep} k={k} -> {digits} {cols}: {ms:.4f} ms", flush=True)
            Z.plan_set(k, current[k][0], current[k][1], kind=Z.PLAN_BATCH)
            Z.lib().zkgpu_drop_tables(0)
    if verbose:
        print(f"BEST mix={mix}: {best_ms:.4f} ms", flush=True)
    if out:
        existing = {}
        if os.path.exists(out):
            for line in open(out):
                f = line.split()
                if len(f) >= 9 and not line.startswith("#"):
                    existing[int(f[0])] = line.rstrip("\n")
        for k, (digits, cols) in current.items():
            d = list(digits) + [0] * (3 - len(digits))
            c = list(cols) + [0] * (3 - len(cols))
            existing[k] = f"{k} {len(digits)} {d[0]} {d[1]} {d[2]} {c[0]} {c[1]} {c[2]} {best_ms:.5f}"
        with open(out, "w") as f:
            f.write(f"# log_n npass d0 d1 d2 c0 c1 c2 ms  (autotuned on gfx950 for batches, mix {mix}: "
                    f"whole-batch ms)\n")
            for k in sorted(existing):
                f.write(existing[k] + "\n")
    return current, best_ms


def _parse_ks(s):
    ks = []
    for part in s.split(","):
        if "-" in part:
            a, b = part.split("-")
            ks.extend(range(int(a), int(b) + 1))
        else:
            ks.append(int(part))
    return ks


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--k", default="12-24")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None)
    ap.add_argument("--kind", choices=["single", "batch"], default="single")
    ap.add_argument("--max-candidates", type=int, default=None)
    ap.add_argument("--mix", default=None, help="tune batch plans on a mixed workload (mixed64, serve)")
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--wide", action="store_true", help="let each pass choose its own tile width")
    a = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    t = time.time()
    if a.mix:
        tune_batch_mix(a.mix, reps=a.reps, sweeps=a.sweeps, out=a.out, wide=a.wide)
    else:
        tune(_parse_ks(a.k), reps=a.reps, out=a.out, max_candidates=a.max_candidates, kind=a.kind, wide=a.wide)
    print(f"autotune took {time.time() - t:.1f} s")
