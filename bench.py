#!/usr/bin/env python3
"""Spliced-DP benchmark: batched extension DP, 65,536 protein/DNA pairs on one MI355X.

Workload: protein lengths log-uniform 50-1000 aa drawn from BLOSUM62 background frequencies,
reverse-translated, 10-30% substitutions, 0-3 GT..AG introns (log-normal, median 300 nt,
max 5 kb), 0-2 frameshifts, 200-nt random flanks; one right-extension request per pair over
the whole target (int16 saturating scores, BLOSUM62, library default penalties).

A step is one pass of the hot path over the whole batch with the raw inputs already resident
in HBM: device-side preparation (codon translation, splice signals) plus the DP of every pair.
GCUPS counts evaluated cells, (rows evaluated) x al per pair, rows starting at 2 and ending at
the x-drop row; the reference's nl x al window cells are reported beside it.

Output: one JSON line (rank 0).
"""
import argparse
import concurrent.futures as cf
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12   # packed-int16 VALU ops/s of MI355X (256 CUs, 2.4 GHz)
OPS_PER_CELL = 32                               # algorithmic int ops per extension cell (31 + row max)


def cpu_baseline(pairs, kw, seconds=15.0, max_threads=16):
    """the reference CPU path (nasw-sse.c via oracle/_ref when built, else the oracle port) on a
    bounded sample of the same workload, threaded like kt_for"""
    import nsffi
    if nsffi.have_ref():
        lib, kind = nsffi.Reference(), "reference"
        fn = lambda d, p: lib.run("ns_global_gs16b", d, p, lib.opt(**kw))
    else:
        lib, kind = nsffi.Oracle(), "port"
        fn = lambda d, p: lib.ext(d, p, lib.opt(**kw))[0]
    try:
        threads = min(len(os.sched_getaffinity(0)), max_threads)
    except AttributeError:
        threads = min(os.cpu_count() or 1, max_threads)
    # chunks of pairs until about `seconds` of wall time have been spent
    n, t0 = 0, time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        while n < len(pairs) and time.perf_counter() - t0 < seconds:
            chunk = pairs[n:n + 4 * threads]
            list(ex.map(lambda x: fn(*x), chunk))
            n += len(chunk)
    dt = time.perf_counter() - t0
    sample = pairs[:n]
    win = sum(len(d) * len(p) for d, p in sample)
    return {"value": win / dt / 1e9, "unit": "GCUPS (nl*al window cells)", "cores": threads, "kind": kind,
            "sample": "%d of the pairs, %.1f s wall" % (n, dt)}, n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; the DP problems are "
              "independent, so N GPUs would shard pairs by cost (see DESIGN.md) - run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import nsffi
    from protlocus_amd import synth
    t0 = time.time()
    pairs = synth.extension_batch(args.pairs, seed=args.seed)
    print("[bench] generated %d pairs in %.1f s" % (len(pairs), time.time() - t0), file=sys.stderr, flush=True)
    prod = nsffi.Product()
    if not prod.lib.ns_gpu_init(0):
        raise SystemExit("GPU not usable: " + prod.error())
    kw = dict(flag=nsffi.NS_F_EXT_RIGHT)
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    plan = prod.ext_plan(ns, ns_off, nl, as_, as_off, al, prod.opt(**kw))
    for _ in range(args.warmup):
        plan.run()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    prod.lib.nsh_device_synchronize()
    dp_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(sync=False)
        plan.prod.lib.nsh_ext_plan_sync(plan.p)
        dp_ms.append(plan.ms(1))
    prod.lib.nsh_device_synchronize()
    wall = time.perf_counter() - t0
    res = plan.results()
    cells = plan.cells(1)          # evaluated cells of the last run
    win = plan.cells(0)
    plan.close()
    ms = wall / args.steps * 1e3
    gcups = cells / (ms * 1e-3) / 1e9
    dp_avg = sum(dp_ms) / len(dp_ms)
    achieved = cells * OPS_PER_CELL / (dp_avg * 1e-3) / 1e12
    out = {
        "metric": "spliced-DP GCUPS (batched extension DP)",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int16", "data": "synthetic",
        "config": {"workload": "batched extension DP, %d pairs, protein 50-1000 aa log-uniform, "
                               "right extension over the full target" % args.pairs,
                   "pairs": args.pairs, "seed": args.seed, "cell_definition": "(rows evaluated) x al"},
        "proteins_per_s": round(args.pairs / (ms * 1e-3), 1),
        "window_gcups": round(win / (ms * 1e-3) / 1e9, 2),
        "dp_kernel_ms": round(dp_avg, 3),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                     "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                     "ops_per_cell": OPS_PER_CELL},
    }
    print("[bench] gpu: %.2f GCUPS, %.1f ms/step, dp %.1f ms" % (gcups, ms, dp_avg), file=sys.stderr, flush=True)
    if not args.no_cpu:
        cb, n = cpu_baseline(pairs, kw, seconds=args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_window_vs_cpu"] = round(out["window_gcups"] / cb["value"], 1) if cb["value"] else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
