#!/usr/bin/env python3
"""Spliced-DP benchmark on one MI355X.

Headline (value): batched extension DP, 65,536 protein/DNA pairs. Protein lengths log-uniform
50-1000 aa drawn from BLOSUM62 background frequencies, reverse-translated, 10-30%
substitutions, 0-3 GT..AG introns (log-normal, median 300 nt, max 5 kb), 0-2 frameshifts,
200-nt random flanks; one right-extension request per pair over the whole target (int16
saturating scores, BLOSUM62, library default penalties). A step is one pass of the hot path
over the whole batch with the raw inputs already resident in HBM: device-side preparation
(codon translation, splice signals) plus the DP of every pair. GCUPS counts evaluated cells,
(rows evaluated) x al per pair, rows starting at 2 and ending at the x-drop row; the
reference's nl x al window cells are reported beside it.

Beside it, in the same JSON line:
  * prescreen: the learned chain-candidate prescreen trained on the GPU, then the plain and the
    prescreened DP pipelines timed on fresh candidates (recall against the exact DP);
  * hmm: profile-HMM Viterbi scan, 64 profiles (M 100-800) x a 10-Mb genome in 6 frames;
  * profile_search: the profile prescreen search of the same profiles over the same genome;
  * gapless_prefilter: gapless scores and top-K selection of one --gpu-prefilter mini-batch;
  * index: the k-mer block index of a 100-Mb genome, and the anchors of 20,000 proteins against it;
  * map: the whole map path (nsh_map_batch) on the genome-scale workload, 20,000 proteins against
    100 Mb with planted gene models, in proteins per second, beside the reference's own program
    on the host cores, with its records compared;
  * map_prescreen: the learned prescreen on the map path's own candidates (trained on a second
    such workload, labels = which candidates became records), the map path timed plain and
    prescreened, with the record recall.
Output: one JSON line.
"""
import argparse
import concurrent.futures as cf
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
# An extension plan runs its column-count groups on streams of their own, which overlap only on
# distinct hardware queues; HIP's default of 4 queues per process is shared by every stream the
# engines create, so give the process more before HIP starts (the runtime reads it once).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "32")
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

VALU_I32_TOPS = 256 * 64 * 2.4e9 / 1e12          # int32 VALU ops/s of MI355X (256 CUs, 2.4 GHz)
VALU_PEAK_TOPS = 2 * VALU_I32_TOPS                # packed int16 (two 16-bit ops per lane)
MFMA_BF16_TFLOPS = 2500.0                         # dense BF16 MFMA
OPS_PER_CELL = 32                                 # extension DP: 31 recurrence ops + row max
HMM_OPS_PER_CELL = 17                             # Viterbi: 8 adds, 4 max, 3 clamps, 2 running-best max
MLP_FLOP_PER_SAMPLE = 2 * (64 * 256 + 256 * 256 + 256)
# HBM bytes of one DP phase (both column-group kernels) of the default workload (65,536 pairs,
# seed 1), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of one step (KiB as read, x 1024). The
# counters are scaled per access path by what they read for reads of known size: FETCH_SIZE reads
# 1.0 x a 1-GiB read through the scalar data cache (nsh_diag_scalar_read, the path of the row
# records) and 0.5 x a 1-GiB runtime device copy (the vector path); WRITE_SIZE reads 1.0 x. The
# scalar-path bytes are the row records the DP reads (4 bytes per evaluated row, the rows read
# ahead, a half 64-byte line per problem); the rest of FETCH_SIZE is taken as vector-path reads.
# The summary and the commands are in profiles/ext_dp_kernel_65536pairs.md
EXT_DP_FETCH_RAW_BYTES = 585703.8 * 1024
EXT_DP_WRITE_BYTES = 74018.2 * 1024
FETCH_SCALE_SCALAR, FETCH_SCALE_VECTOR = 1.0, 2.0


def ext_dp_traffic(rows_done, n_problems):
    """HBM bytes per DP phase of the default workload (see the constants above)"""
    scalar = min(EXT_DP_FETCH_RAW_BYTES, 4.0 * (rows_done + 4 * n_problems) + 32.0 * n_problems)
    return (FETCH_SCALE_SCALAR * scalar + FETCH_SCALE_VECTOR * (EXT_DP_FETCH_RAW_BYTES - scalar) +
            EXT_DP_WRITE_BYTES)


def threads():
    """host threads of the CPU baselines: the box's CPU share for one GPU (OMP_NUM_THREADS, set to 16
    there), else every core this process may run on"""
    try:
        n_aff = len(os.sched_getaffinity(0))
    except AttributeError:
        n_aff = os.cpu_count() or 1
    try:
        return max(1, min(n_aff, int(os.environ["OMP_NUM_THREADS"])))
    except (KeyError, ValueError):
        return n_aff


def host_cpu():
    """the host CPU's model name and the cores this process may run on"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n_aff = len(os.sched_getaffinity(0))
    except AttributeError:
        n_aff = os.cpu_count() or 1
    return {"model": model, "affinity_cores": n_aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(pairs, kw, seconds=15.0, packed=None):
    """the reference CPU path on the same workload: nasw-sse.c's ns_global_gs16b compiled from the
    reference sources (oracle/_ref), driven from C threads that take requests from a shared counter
    as kt_for does, over every pair unless `seconds` runs out first; without oracle/_ref the oracle
    port on a bounded sample from Python threads. Returns (baseline object, reference results by
    pair index or None)."""
    import numpy as np
    import nsffi
    nt = threads()
    if nsffi.have_ref():
        ref = nsffi.Reference()
        ns, ns_off, nl, as_, as_off, al = packed
        out = np.full(3 * len(pairs), -1, dtype=np.int32)
        wall = C.c_double()
        opt = ref.opt(**kw)
        ref.lib.or_ref_ext_batch.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_void_p,
                                             C.c_void_p]
        ref.lib.or_ref_ext_batch.restype = C.c_int32
        ns_off = np.ascontiguousarray(ns_off, dtype=np.int64)
        as_off = np.ascontiguousarray(as_off, dtype=np.int64)
        nl = np.ascontiguousarray(nl, dtype=np.int32)
        al = np.ascontiguousarray(al, dtype=np.int32)
        done = ref.lib.or_ref_ext_batch(len(pairs), ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data,
                                        al.ctypes.data, None, C.byref(opt), nt, seconds, out.ctypes.data,
                                        C.byref(wall))
        res = out.reshape(-1, 3)
        got = res[:, 0] >= 0
        win = float((nl[got].astype(np.float64) * al[got]).sum())
        base = {"value": win / wall.value / 1e9, "unit": "GCUPS (nl*al window cells)", "cores": nt,
                "kind": "reference", "cpu": host_cpu(),
                "sample": "%d of the %d pairs (%s), %.1f s wall, nasw-sse.c ns_global_gs16b from %d C threads"
                          % (done, len(pairs), "all" if done == len(pairs) else "time-bounded", wall.value, nt)}
        return base, (res if done == len(pairs) else None)
    lib, kind = nsffi.Oracle(), "port"
    fn = lambda d, p: lib.ext(d, p, lib.opt(**kw))[0]
    n, t0 = 0, time.perf_counter()
    with cf.ThreadPoolExecutor(nt) as ex:
        while n < len(pairs) and time.perf_counter() - t0 < seconds:
            chunk = pairs[n:n + 4 * nt]
            list(ex.map(lambda x: fn(*x), chunk))
            n += len(chunk)
    dt = time.perf_counter() - t0
    win = sum(len(d) * len(p) for d, p in pairs[:n])
    return {"value": win / dt / 1e9, "unit": "GCUPS (nl*al window cells)", "cores": nt, "kind": kind,
            "cpu": host_cpu(), "sample": "%d of the pairs, %.1f s wall (oracle port, Python threads)" % (n, dt)}, None


def _drivers():
    lib = C.CDLL(os.path.join(ROOT, "protlocus_amd", "libnsh_drivers.so"))
    for f in (lib.nshd_ext_batch,):
        f.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p,
                      C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    lib.nshd_ext_single.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p,
                                    C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.nshd_gpu_submit.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    for f in (lib.nshd_ext_batch, lib.nshd_ext_single, lib.nshd_gpu_submit):
        f.restype = C.c_int32
    return lib


def bench_dropin(prod, pairs, packed, kw, ref_res, log, n_single=16384, n_global=16384, cpu=True):
    """the drop-in path: the reference's entry points driven from C threads the way align.c calls
    them (extension batches per query, single extension calls, ns_gpu_submit/wait of gap DPs),
    against the resident-plan headline; results against the reference's"""
    import numpy as np
    import nsffi
    from protlocus_amd import synth
    drv = _drivers()
    nt = threads()
    ns, ns_off, nl, as_, as_off, al = packed
    ns_off = np.ascontiguousarray(ns_off, dtype=np.int64)
    as_off = np.ascontiguousarray(as_off, dtype=np.int64)
    nl = np.ascontiguousarray(nl, dtype=np.int32)
    al = np.ascontiguousarray(al, dtype=np.int32)
    opt = prod.opt(**kw)
    win = float((nl.astype(np.float64) * al).sum())
    out = {"threads": nt, "workload": "the %d headline pairs" % len(pairs)}
    res = np.zeros(3 * len(pairs), dtype=np.int32)
    wall = C.c_double()
    for rep in range(2):   # the first pass grows the plan pool
        ok = drv.nshd_ext_batch(len(pairs), ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data,
                                al.ctypes.data, None, C.byref(opt), nt, 8, res.ctypes.data, C.byref(wall))
    r = res.reshape(-1, 3)
    same = None if ref_res is None else int((r == ref_res).all(axis=1).sum())
    out["ext_batch_api"] = {"pattern": "per group of 8 requests (one query's candidates): ns_ext_batch_init, add, flush, "
                                       "ready/wait, get, destroy; %d C threads" % nt,
                            "ok": bool(ok), "wall_s": round(wall.value, 3), "requests_per_s": round(len(pairs) / wall.value, 1),
                            "window_gcups": round(win / wall.value / 1e9, 2),
                            "identical_to_reference": None if same is None else "%d of %d" % (same, len(pairs))}
    log("[bench] drop-in ext batch API: %.1f window GCUPS, %.0f req/s" % (win / wall.value / 1e9, len(pairs) / wall.value))
    m = min(n_single, len(pairs))
    res1 = np.zeros(3 * m, dtype=np.int32)
    ok = drv.nshd_ext_single(m, ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data, al.ctypes.data, None,
                             C.byref(opt), nt, res1.ctypes.data, C.byref(wall))
    r1 = res1.reshape(-1, 3)
    same1 = None if ref_res is None else int((r1 == ref_res[:m]).all(axis=1).sum())
    win1 = float((nl[:m].astype(np.float64) * al[:m]).sum())
    out["ext_single_calls"] = {"pattern": "one ns_global_gs16b per request from %d C threads" % nt, "requests": m,
                               "ok": bool(ok), "wall_s": round(wall.value, 3), "requests_per_s": round(m / wall.value, 1),
                               "window_gcups": round(win1 / wall.value / 1e9, 2),
                               "identical_to_reference": None if same1 is None else "%d of %d" % (same1, m)}
    log("[bench] drop-in single ext calls: %.1f window GCUPS" % (win1 / wall.value / 1e9))
    # global DP with CIGAR through ns_gpu_submit / ns_gpu_wait
    gp = synth.extension_batch(n_global, seed=5, min_len=20, max_len=400)
    gns, gns_off, gnl, gas, gas_off, gal = synth.pack_batch(gp)
    gns_off = np.ascontiguousarray(gns_off, dtype=np.int64)
    gas_off = np.ascontiguousarray(gas_off, dtype=np.int64)
    gnl = np.ascontiguousarray(gnl, dtype=np.int32)
    gal = np.ascontiguousarray(gal, dtype=np.int32)
    gopt = prod.opt(flag=nsffi.NS_F_CIGAR)
    cells = float(((gnl.astype(np.float64) - 2).clip(0) * gal).sum())
    gres = np.zeros(3 * n_global, dtype=np.int32)
    L = prod.lib
    L.ns_gpu_kernel_seconds.restype = C.c_double
    L.ns_gpu_d2h_seconds.restype = C.c_double
    for rep in range(2):
        k0, d0 = L.ns_gpu_kernel_seconds(), L.ns_gpu_d2h_seconds()
        ok = drv.nshd_gpu_submit(n_global, gns, gns_off.ctypes.data, gnl.ctypes.data, gas, gas_off.ctypes.data,
                                 gal.ctypes.data, C.byref(gopt), nt, 64, gres.ctypes.data, C.byref(wall))
        k1, d1 = L.ns_gpu_kernel_seconds(), L.ns_gpu_d2h_seconds()
    gr = gres.reshape(-1, 3)
    gcups = cells / wall.value / 1e9
    g = {"workload": "%d global CIGAR problems (extension-batch generator, protein 20-400 aa, 200-nt flanks), "
                     "cells = (nl - 2) x al" % n_global,
         "pattern": "per group of 64: ns_gpu_submit of each, then ns_gpu_wait of each; %d C threads" % nt,
         "ok": bool(ok), "wall_s": round(wall.value, 3), "cells": cells, "gcups": round(gcups, 2),
         "engine_batch_s": round(k1 - k0, 3), "cigar_d2h_s": round(d1 - d0, 4),
         "roofline": {"bound": "valu", "achieved": round(gcups * 31 / 1e3, 3), "peak": round(VALU_I32_TOPS, 2),
                      "unit": "TOP/s", "frac": round(gcups * 31 / 1e3 / VALU_I32_TOPS, 4), "ops_per_cell": 31}}
    if cpu and nsffi.have_ref():
        ref = nsffi.Reference()
        L2 = ref.lib
        L2.or_ref_global_batch.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        ms = min(n_global, 4096)
        rr = np.zeros(2 * ms, dtype=np.int32)
        ropt = ref.opt(flag=nsffi.NS_F_CIGAR)
        L2.or_ref_global_batch(ms, gns, gns_off.ctypes.data, gnl.ctypes.data, gas, gas_off.ctypes.data, gal.ctypes.data,
                               C.byref(ropt), nt, rr.ctypes.data, C.byref(wall))
        rr = rr.reshape(-1, 2)
        same_g = int((rr == gr[:ms, :2]).all(axis=1).sum())
        ccells = float(((gnl[:ms].astype(np.float64) - 2).clip(0) * gal[:ms]).sum())
        g["identical_to_reference"] = "%d of %d (score and CIGAR length)" % (same_g, ms)
        g["cpu_baseline"] = {"value": round(ccells / wall.value / 1e9, 3), "unit": "GCUPS", "cores": nt, "kind": "reference",
                             "cpu": host_cpu(), "sample": "first %d problems, %.2f s, ns_global_gs16b with CIGAR" % (ms, wall.value)}
    # the same problems as one engine batch (nsh_global_batch), as a pipeline hands them over
    L.nsh_global_batch.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p,
                                   C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.POINTER(C.c_uint32))]
    sc = np.zeros(n_global, dtype=np.int32)
    nc = np.zeros(n_global, dtype=np.int32)
    best = 1e30
    for rep in range(3):
        cig = C.POINTER(C.c_uint32)()
        t0 = time.perf_counter()
        okb = L.nsh_global_batch(n_global, gns, gns_off.ctypes.data, gnl.ctypes.data, gas, gas_off.ctypes.data,
                                 gal.ctypes.data, None, C.byref(gopt), 0, sc.ctypes.data, nc.ctypes.data, C.byref(cig))
        best = min(best, time.perf_counter() - t0)
        prod.free(C.cast(cig, C.c_void_p))
    bg = cells / best / 1e9
    g["batch"] = {"entry": "nsh_global_batch, all problems in one engine batch (host inputs, CIGARs back on the host)",
                  "ok": bool(okb), "wall_s": round(best, 4), "gcups": round(bg, 2),
                  "same_as_submit_path": bool((sc == gr[:, 0]).all() and (nc == gr[:, 1]).all()),
                  "roofline": {"bound": "valu", "achieved": round(bg * 31 / 1e3, 3), "peak": round(VALU_I32_TOPS, 2),
                               "unit": "TOP/s", "frac": round(bg * 31 / 1e3 / VALU_I32_TOPS, 4), "ops_per_cell": 31}}
    out["global_cigar"] = g
    log("[bench] drop-in global CIGAR: %.1f GCUPS; one batch %.1f GCUPS" % (gcups, bg))
    return out


def bench_hmm(seconds_cpu=4.0, log=None):
    """64 profiles x a 10-Mb genome in 6 frames: Viterbi cells/s on the GPU (kernel time) and the
    reference's mp_hmm_dp_score (or the oracle) on a bounded sample of the same regions"""
    import numpy as np
    import nsffi
    from protlocus_amd import capi, synth
    t0 = time.time()
    profs, nt, off, ln = synth.profile_scan()
    rows = (ln // 3).astype(np.int64)
    if log:
        log("[bench] hmm workload: %d profiles, %d regions, %.1f Mb in %.1f s" % (len(profs), len(ln), len(nt) / 1e6,
                                                                              time.time() - t0))
    cells = 0.0
    kern_ms = 0.0
    t0 = time.perf_counter()
    best = []
    for em, tr in profs:
        e, t = synth.hmm_tables(em, tr)
        h = capi.HMM(e, t)
        try:
            s = h.score(nt, off, ln)
            kern_ms += h.last_ms()
        finally:
            h.close()
        cells += float(rows.sum()) * len(e)
        best.append(int(s.max()))
    wall = time.perf_counter() - t0
    gcups = cells / (kern_ms * 1e-3) / 1e9
    out = {"workload": "64 profile HMMs (M 100-800, log-uniform) x 10 Mb genome (GC 0.45, 200 planted hits), "
                       "3 kb windows in 6 frames", "profiles": len(profs), "regions": int(len(ln)),
           "cells": cells, "kernel_ms": round(kern_ms, 2), "gcups": round(gcups, 1),
           "wall_s_with_uploads": round(wall, 2),
           "roofline": {"bound": "valu", "achieved": round(gcups * HMM_OPS_PER_CELL / 1e3, 3),
                        "peak": round(VALU_I32_TOPS, 2), "unit": "TOP/s",
                        "frac": round(gcups * HMM_OPS_PER_CELL / 1e3 / VALU_I32_TOPS, 4),
                        "ops_per_cell": HMM_OPS_PER_CELL}}
    # CPU: one profile (the first) over a sample of regions, threaded
    e, t = synth.hmm_tables(*profs[0])
    text = synth.hmm_text(*profs[0])
    if nsffi.have_ref():
        ref = nsffi.Reference()
        with tempfile.NamedTemporaryFile("w", suffix=".hmm", delete=False) as f:
            f.write(text)
        fp = ref.lib.mp_hmm_open(f.name.encode())
        hh = ref.lib.mp_hmm_read1(fp)
        ref.lib.mp_hmm_close(fp)
        os.unlink(f.name)
        fn = lambda k: ref.lib.mp_hmm_dp_score(hh, nt[off[k]:off[k] + ln[k]], int(ln[k]))
        kind = "reference"
    else:
        ora = nsffi.Oracle()
        hh = nsffi.OrHmm()
        ora.lib.or_hmm_parse(text.encode(), C.byref(hh), None)
        fn = lambda k: ora.lib.or_hmm_viterbi(C.byref(hh), nt[off[k]:off[k] + ln[k]], int(ln[k]))
        kind = "port"
    ntr = threads()
    n, t0 = 0, time.perf_counter()
    with cf.ThreadPoolExecutor(ntr) as ex:
        while n < len(ln) and time.perf_counter() - t0 < seconds_cpu:
            ks = list(range(n, min(len(ln), n + 8 * ntr)))
            list(ex.map(fn, ks))
            n += len(ks)
    dt = time.perf_counter() - t0
    cpu_cells = float(rows[:n].sum()) * len(e)
    out["cpu_baseline"] = {"value": round(cpu_cells / dt / 1e9, 3), "unit": "GCUPS", "cores": ntr, "kind": kind,
                           "sample": "profile 0 (M=%d) x %d regions, %.1f s wall" % (len(e), n, dt)}
    return out


def ps_kmer_thr(kmer, sens):
    """mp_ps_kmer_thr (prescreen.c:63-71) in its float32 arithmetic"""
    import numpy as np
    f = np.float32
    a, b = {5: (108.8, 4.7), 7: (149.15, 6.85)}.get(kmer, (134.35, 6.15))
    t = f(a) - f(sens) * f(b)
    return int(np.floor(f(int(t)) / f(4.0) + f(0.5)))


def bench_profile_search(cpu=True, seconds_cpu=10.0, log=None):
    """the profile prescreen search (mp_prescreen_build, then ps_query_worker per profile) over the
    HMM workload's 10-Mb genome and 64 profiles: the GPU's index build and search times and
    candidates, and the reference's own search timed on a bounded sample of the profiles (threads
    over profiles), whose candidates are compared with the GPU's"""
    import numpy as np
    import nsffi
    from protlocus_amd import capi, synth
    kmer, min_orf, sens, max_ev = 6, 30, 7.5, 1e-3
    thr = ps_kmer_thr(kmer, sens)
    profs, g = synth.profile_genome()
    pssms = [synth.hmm_pssm(em) for em, _ in profs]
    t0 = time.perf_counter()
    ps = capi.ProfileSearch(g, [0], [len(g)], kmer=kmer, min_orf=min_orf)
    build_s = time.perf_counter() - t0
    try:
        sizes = (ps.size(0), ps.size(1), ps.size(2))
        ps.search_batch(pssms, thr, max_evalue=max_ev)   # warm-up (workspace growth)
        t0 = time.perf_counter()
        res = ps.search_batch(pssms, thr, max_evalue=max_ev)
        batch_s = time.perf_counter() - t0
        counts = ps.stats()
        t0 = time.perf_counter()
        single = [ps.search(P, thr, max_evalue=max_ev) for P in pssms]
        search_s = time.perf_counter() - t0
    finally:
        ps.close()
    assert single == res, "batched and per-profile searches disagree"
    out = {"workload": "64 profile HMMs (M 100-800) x 10 Mb genome (GC 0.45, 200 planted hits), both strands: "
                       "ORF database + spaced 6-mer index, then per profile k-mer hits, double-hit diagonals, "
                       "ungapped and banded gapped scores, E <= 1e-3",
           "orfs": sizes[0], "orf_residues": sizes[1], "index_entries": sizes[2],
           "gpu_build_ms": round(build_s * 1e3, 1),
           "gpu_batch_ms_per_profile": round(batch_s * 1e3 / len(pssms), 3),
           "gpu_single_ms_per_profile": round(search_s * 1e3 / len(pssms), 3),
           "candidates": int(sum(len(r) for r in res)), "search_counts": counts}
    if log:
        log("[bench] profile search: build %.0f ms, %.3f ms per profile batched (%.3f one by one), %d candidates" % (
            build_s * 1e3, batch_s * 1e3 / len(pssms), search_s * 1e3 / len(pssms), out["candidates"]))
    if cpu and nsffi.have_ref():
        ref = nsffi.Reference()
        t0 = time.perf_counter()
        hnd = ref.ps_build(g, [0], [len(g)], kmer=kmer, min_orf=min_orf)
        cpu_build = time.perf_counter() - t0
        hmms = [ref.hmm(synth.hmm_text(em, tr)) for em, tr in profs]

        def one(k):
            return ref.ps_query(hnd, hmms[k], kmer=kmer, sens=sens, max_evalue=max_ev)
        ntr = threads()
        done, same, t0 = [], 0, time.perf_counter()
        with cf.ThreadPoolExecutor(ntr) as ex:
            while len(done) < len(profs) and time.perf_counter() - t0 < seconds_cpu:
                ks = list(range(len(done), min(len(profs), len(done) + ntr)))
                done += list(ex.map(one, ks))
        dt = time.perf_counter() - t0
        same = sum(a == b for a, b in zip(done, res))
        ref.ps_free(hnd)
        for h in hmms:
            ref.lib.mp_hmm_destroy(h)
        out["cpu_baseline"] = {"value": round(dt * 1e3 / len(done), 2), "unit": "ms per profile (wall, %d threads)" % ntr,
                               "cores": ntr, "kind": "reference", "build_ms": round(cpu_build * 1e3, 1),
                               "sample": "%d of 64 profiles, %.1f s wall" % (len(done), dt)}
        out["identical_candidates"] = "%d of %d sampled profiles" % (same, len(done))
    return out


def bench_gapless_prefilter(cpu=True, reps=5, log=None):
    """the --gpu-prefilter stage of one mini-batch (mp_gpu_prefilter_batch, map.c:1014-1089): GPU
    gapless scores of every coarse candidate, then the batch calibration, adaptive K and top-K
    selection (hit.c:149-263). Times the fused call, the scoring alone and the selection alone on
    the GPU, and the reference's selection (oracle/_ref, one host thread, as the reference runs it
    on its pipeline thread) on the same scores, whose keep flags must be identical."""
    import numpy as np
    import nsffi
    from protlocus_amd import capi, synth
    b = synth.gapless_minibatch()
    prod = nsffi.Product()
    L = prod.lib
    g = b["genome"]
    pk = (np.append(g, 0) if len(g) & 1 else g)
    pk = (pk[0::2] | (pk[1::2] << 4)).astype(np.uint8).tobytes()
    n, n_q = len(b["vid"]), len(b["qlen"])
    ca = np.zeros(n, dtype=[("target_offset", "<i8"), ("target_length", "<i8"), ("query_offset", "<i8"),
                            ("query_length", "<i4"), ("anchor_offset", "<i4"), ("anchor_count", "<i4"),
                            ("reverse", "<i4")])
    for f in ca.dtype.names:
        ca[f] = b[f]
    an = np.zeros(len(b["anchor_block"]), dtype=[("target_block", "<i8"), ("query_end", "<i4"), ("reserved", "<i4")])
    an["target_block"], an["query_end"] = b["anchor_block"], b["anchor_qend"]
    CA = (nsffi.GaplessCand * n).from_buffer(ca)
    AN = (nsffi.GaplessAnchor * len(an)).from_buffer(an)
    mat = np.frombuffer(bytes((C.c_int8 * 484).in_dll(L, "ns_mat_blosum62")), dtype=np.int8)
    codon = bytes((C.c_uint8 * 64).in_dll(L, "ns_tab_codon"))
    queries = b["queries"].tobytes()
    bw, kmer, xdrop, top_k, mpv = 256, 6, 40, 32, 1
    sa = [np.ascontiguousarray(x, dtype=t) for x, t in ((b["cand_off"], np.int32), (b["qlen"], np.int32),
                                                        (b["chn"], np.int32), (b["vid"], np.uint32), (b["vs"], np.int64))]
    sel_args = tuple(sa)
    scores = np.zeros(n, dtype=np.int32)
    L.ns_gpu_gapless_reset_target()

    def score():
        ok = L.ns_gpu_gapless_score_batch(pk, len(g), queries, len(queries), CA, n, AN, len(an), bw, kmer, xdrop,
                                          mat.ctypes.data, 22, codon, scores.ctypes.data)
        assert ok, prod.error()
    out = {"workload": "one --gpu-prefilter mini-batch: %d queries (50-1000 aa, %d residues), %d coarse candidates "
                       "(2-48 per query, one true locus each, 20%% substitutions), %d anchors, 20 Mb genome in 24 "
                       "contigs; block 256, k-mer 6, X-drop 40, top-K 32, per-contig floor 1" % (
                           n_q, len(queries), n, len(an)),
           "candidates": n}
    for adaptive in (1, 0):
        fused = capi.gapless_prefilter(pk, len(g), queries, CA, AN, bw, kmer, xdrop, mat, codon, *sel_args, top_k,
                                       mpv, adaptive, as_numpy=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            capi.gapless_prefilter(pk, len(g), queries, CA, AN, bw, kmer, xdrop, mat, codon, *sel_args, top_k, mpv,
                                   adaptive, as_numpy=True)
        fused_ms = (time.perf_counter() - t0) * 1e3 / reps
        score()
        t0 = time.perf_counter()
        for _ in range(reps):
            score()
        score_ms = (time.perf_counter() - t0) * 1e3 / reps
        assert np.array_equal(scores, fused[0]), "fused and separate gapless scores disagree"
        capi.gapless_select(sa[0], sa[1], scores, *sa[2:], top_k, mpv, adaptive, as_numpy=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            sel = capi.gapless_select(sa[0], sa[1], scores, *sa[2:], top_k, mpv, adaptive, as_numpy=True)
        sel_ms = (time.perf_counter() - t0) * 1e3 / reps
        assert np.array_equal(sel[0], fused[1]) and np.array_equal(sel[1], fused[2])
        key = "adaptive" if adaptive else "fixed_k"
        r = {"fused_ms": round(fused_ms, 3), "score_ms": round(score_ms, 3), "select_ms": round(sel_ms, 3),
             "kept": int(sel[0].sum()), "calibration": {"noise": sel[2][0], "n_query": sel[2][1],
                                                      "dup_frac": round(sel[2][2], 4)}}
        if cpu and nsffi.have_ref():
            ref = nsffi.Reference()
            t0 = time.perf_counter()
            want = ref.gapless_select(b["cand_off"], b["qlen"], scores, b["chn"], b["vid"], b["vs"], top_k, mpv,
                                      adaptive)
            cpu_ms = (time.perf_counter() - t0) * 1e3
            r["cpu_baseline"] = {"value": round(cpu_ms, 3), "unit": "ms per mini-batch (selection only)", "cores": 1,
                                 "kind": "reference", "sample": "the whole mini-batch, one host thread"}
            r["identical_to_reference"] = bool(want[0] == sel[0].tolist() and want[1] == sel[1].tolist() and
                                               want[2] == sel[2])
        out[key] = r
        if log:
            log("[bench] gapless prefilter (%s): fused %.2f ms, score %.2f ms, select %.3f ms, kept %d of %d" % (
                key, fused_ms, score_ms, sel_ms, r["kept"], n))
    return out


def bench_index(cpu=True, genome_len=100_000_000, n_ctg=24, log=None):
    """the k-mer block index of a 100-Mb genome (mp_idx_build's mp_idx_boff, mp_sketch_nt4 per
    contig strand and build_bidx; index options kmer 6, mod_bit 1, bbit 8, min_aa_len 30): the GPU
    build (genome upload included), and the reference's code (oracle/_ref) on one host thread, whose
    bucket starts and entries must be identical"""
    import numpy as np
    import nsffi
    from protlocus_amd import capi
    rng = np.random.default_rng(17)
    g = rng.choice(4, size=genome_len, p=[0.275, 0.225, 0.225, 0.275]).astype(np.uint8)
    for s0 in rng.integers(0, genome_len - 1000, size=2000):
        g[s0:s0 + int(rng.integers(1, 500))] = 4
    cuts = np.sort(rng.choice(np.arange(10_000, genome_len - 10_000), n_ctg - 1, replace=False))
    off = np.concatenate([[0], cuts]).astype(np.int64)
    ln = np.diff(np.concatenate([off, [genome_len]])).astype(np.int64)
    nsffi.Product()   # tables
    capi.index_build(g[:100000], [0], [100000])   # warm-up
    L = capi.lib()
    t0 = time.perf_counter()
    h = L.nsh_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, 30, 6, 1, 8)
    gpu_s = time.perf_counter() - t0
    assert h, capi.error()
    L.nsh_index_destroy(h)
    pk = capi.pack_nt4(g)   # the caller's genome as mp_ntseq_read stores it (4 bits per base)
    t0 = time.perf_counter()
    h = L.nsh_index_build_nt4(len(off), pk.ctypes.data, off.ctypes.data, ln.ctypes.data, 30, 6, 1, 8)
    nt4_s = time.perf_counter() - t0
    assert h, capi.error()
    L.nsh_index_destroy(h)
    t0 = time.perf_counter()
    ki, kb, bo = capi.index_build(g, off, ln)
    copy_s = time.perf_counter() - t0
    out = {"workload": "k-mer block index of a %d-Mb genome in %d contigs (GC 0.45, 2,000 N runs), both strands: "
                       "6-mers of the 13-letter alphabet over ORFs >= 30 codons, mod_bit 1, 256-base blocks" % (
                           genome_len // 1_000_000, n_ctg),
           "entries": int(len(kb)), "buckets": int(len(ki)), "blocks": int(bo[-1]),
           "gpu_ms": round(gpu_s * 1e3, 1), "gpu_mbases_per_s": round(genome_len / gpu_s / 1e6, 1),
           "gpu_nt4_ms": round(nt4_s * 1e3, 1), "gpu_with_copy_out_ms": round(copy_s * 1e3, 1)}
    if log:
        log("[bench] index: %d entries, GPU %.1f ms (from the 4-bit genome %.1f ms)" % (len(kb), gpu_s * 1e3,
                                                                                      nt4_s * 1e3))
    if cpu and nsffi.have_ref():
        ref = nsffi.Reference()
        t0 = time.perf_counter()
        wki, wkb, wbo = ref.index(g, off, ln)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cpu_s * 1e3, 1), "unit": "ms per index (reference code, one host thread)",
                               "cores": 1, "kind": "reference", "sample": "the whole genome"}
        out["identical_to_reference"] = bool(np.array_equal(ki, wki) and np.array_equal(kb, wkb) and
                                             np.array_equal(bo, wbo))
        if log:
            log("[bench] index: reference %.1f ms, identical %s" % (cpu_s * 1e3, out["identical_to_reference"]))
    # the anchors of 20,000 proteins (50-1000 aa): half translated genome windows with 20% substitutions
    from protlocus_amd import synth
    code1 = np.frombuffer(synth.CODE1.encode(), dtype=np.uint8)
    qs = []
    for i in range(20000):
        L = int(np.exp(rng.uniform(np.log(50), np.log(1000))))
        if i % 2 == 0:
            s0 = int(rng.integers(0, genome_len - 3 * L))
            w = g[s0:s0 + 3 * L].astype(np.int64)
            bad = (w.reshape(L, 3) > 3).any(axis=1)
            idx = np.minimum(w[0::3], 3) << 4 | np.minimum(w[1::3], 3) << 2 | np.minimum(w[2::3], 3)
            p = code1[idx].copy()
            p[bad] = ord("X")
            sub = rng.random(L) < 0.2
            p[sub] = np.frombuffer(synth.rand_protein(rng, int(sub.sum())), dtype=np.uint8)
            qs.append(p.tobytes())
        else:
            qs.append(synth.rand_protein(rng, L))
    ix = capi.Index(g, off, ln)
    try:
        ix.anchors(qs[:100])
        t0 = time.perf_counter()
        got = ix.anchors(qs)
        an_s = time.perf_counter() - t0
        ix.chain(qs[:100])
        t0 = time.perf_counter()
        chains = ix.chain(qs)
        ch_s = time.perf_counter() - t0
    finally:
        ix.close()
    n_anchor = int(sum(len(a) for a in got))
    out["anchors"] = {"queries": len(qs), "residues": int(sum(len(q) for q in qs)), "anchors": n_anchor,
                      "gpu_ms": round(an_s * 1e3, 1)}
    if log:
        log("[bench] anchors: %d queries, %d anchors, GPU %.1f ms" % (len(qs), n_anchor, an_s * 1e3))
    if cpu and nsffi.have_ref():
        ref = nsffi.Reference()
        L = ref.lib
        h = L.or_ref_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, 30, 6, 1, 8)
        try:
            same, done, t0 = 0, 0, time.perf_counter()
            buf = np.zeros(1 << 22, np.uint64)
            for k, q in enumerate(qs):
                n = L.or_ref_anchors(h, q, len(q), 20000, len(buf), buf.ctypes.data)
                same += n == len(got[k]) and np.array_equal(buf[:n], got[k])
                done += 1
                if time.perf_counter() - t0 > 10.0:
                    break
            cpu_s = time.perf_counter() - t0
        finally:
            L.or_ref_index_free(h)
        out["anchors"]["cpu_baseline"] = {"value": round(cpu_s * 1e3 / done * len(qs), 1),
                                          "unit": "ms per 20,000 queries (reference code, one host thread, scaled)",
                                          "cores": 1, "kind": "reference",
                                          "sample": "the first %d queries, %.1f s" % (done, cpu_s)}
        out["anchors"]["identical_to_reference"] = "%d of %d sampled queries" % (same, done)
    out["chain"] = bench_chain(got, cpu=cpu, log=log)
    out["index_chain"] = bench_index_chain(g, off, ln, qs, chains, ch_s, cpu=cpu, log=log)
    return out


def bench_map(cpu=True, log=None, n_prot=20000, batch=20000, sample=2000):
    """the whole map path on the genome-scale workload: 20,000 proteins (log-normal lengths, median
    400 aa, max 3000) with one planted gene model each in a 100-Mb genome (24 contigs, GC 0.42, 5% in
    tandem repeats; synth.genome_with_genes seed 21). The index is built once (timed apart); then
    nsh_map_batch maps the proteins (coarse chains, refinement, align stage, records) and the proteins
    per second are over that call. The reference's own program (oracle/_ref/protlocus, its CPU build)
    maps a sample of the same proteins against the same genome at threads() threads; its mapping time
    runs from its index build's last log line to its last "mapped" line, and the PAF records of the
    sampled proteins must be identical to the GPU path's"""
    import re
    import subprocess
    import numpy as np
    from protlocus_amd import capi, synth
    t0 = time.perf_counter()
    cn, cs, pn, ps, scale = synth.genome_with_genes(n_prot=n_prot)
    gen_s = time.perf_counter() - t0
    capi.lib()
    t0 = time.perf_counter()
    mp = capi.Mapper(cn, cs)
    index_s = time.perf_counter() - t0
    try:
        mp.map(pn[:32], ps[:32])   # warm-up: device pools, streams, code objects
        recs, stats = [], np.zeros(24, np.int64)
        t0 = time.perf_counter()
        for b0 in range(0, n_prot, batch):
            r, st = mp.map(pn[b0:b0 + batch], ps[b0:b0 + batch])
            recs.append(r)
            stats += st
        map_s = time.perf_counter() - t0
    finally:
        mp.close()
    paf = b"".join(recs)
    out = {"workload": "%d proteins (log-normal lengths, median 400 aa, max 3000; %d residues) against a 100-Mb "
                       "genome in 24 contigs with one planted gene model per protein (Poisson(6) exons, GT..AG "
                       "introns log-normal, median 1 kb scaled by %.3f to fit, 5-35%% divergence), default options, "
                       "PAF" % (n_prot, sum(len(p) for p in ps), scale),
           "batch": batch, "index_ms": round(index_s * 1e3, 1), "gpu_s": round(map_s, 3),
           "proteins_per_s": round(n_prot / map_s, 1), "records": int(stats[5]), "mapped": int(stats[4]),
           "counts": {"coarse_chains": int(stats[0]), "refine_candidates": int(stats[1]),
                      "refine_pairs": int(stats[2]), "dp_candidates": int(stats[3])}}
    if batch >= n_prot:   # one call: the stage clock is that call's
        prev, stages = 0, {}
        for k, name in enumerate(capi.MAP_STAGES):
            stages[name] = round((stats[8 + k] - prev) / 1e3, 1)
            prev = stats[8 + k]
        out["stage_ms"] = stages
    if log:
        log("[bench] map: %d proteins in %.2f s (%.0f proteins/s), %d records; index %.1f s; synth %.1f s" % (
            n_prot, map_s, n_prot / map_s, stats[5], index_s, gen_s))
    prog = os.path.join(ROOT, "oracle", "_ref", "protlocus")
    if cpu and os.path.exists(prog):
        nth = threads()
        with tempfile.TemporaryDirectory() as d:
            g, p = os.path.join(d, "g.fa"), os.path.join(d, "p.fa")
            synth.write_fasta(g, cn, cs)
            synth.write_fasta(p, pn[:sample], ps[:sample])
            r = subprocess.run([prog, "-t", str(nth), g, p], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                               check=True)
        err = r.stderr.decode()
        t_idx = max(float(x) for x in re.findall(r"\[M::mp_idx_build@([0-9.]+)\*", err))
        t_map = max(float(x) for x in re.findall(r"\[M::worker_pipeline::([0-9.]+)\*[0-9.]+\] mapped", err))
        cpu_s = t_map - t_idx
        names = set(x.encode() for x in pn[:sample])
        mine = b"".join(line + b"\n" for line in paf.split(b"\n") if line and line.split(b"\t", 1)[0] in names)
        same = mine == r.stdout
        out["cpu_baseline"] = {"value": round(sample / cpu_s, 1), "unit": "proteins/s (map stage, index excluded)",
                               "cores": nth, "kind": "reference", "cpu": host_cpu(),
                               "sample": "the first %d proteins against the whole genome: %.2f s of mapping at %d "
                                         "threads" % (sample, cpu_s, nth)}
        out["identical_to_reference"] = bool(same)
        out["gpu_vs_cpu"] = round(out["proteins_per_s"] / out["cpu_baseline"]["value"], 2)
        if not same:
            got = {}
            for line in mine.split(b"\n"):
                if line:
                    got.setdefault(line.split(b"\t", 1)[0], []).append(line)
            exp = {}
            for line in r.stdout.split(b"\n"):
                if line:
                    exp.setdefault(line.split(b"\t", 1)[0], []).append(line)
            bad = [q for q in sorted(names) if got.get(q) != exp.get(q)]
            out["identical_to_reference"] = "%d of %d proteins" % (sample - len(bad), sample)
            out["first_differences"] = [q.decode() for q in bad[:5]]
        if log:
            log("[bench] map: reference %.1f proteins/s at %d threads (%.2f s for %d), identical %s" % (
                sample / cpu_s, nth, cpu_s, sample, out["identical_to_reference"]))
    return out


def bench_index_chain(g, off, ln, qs, chains, gpu_s, cpu=True, log=None):
    """mp_map_coarse's index -> chain slice (map.c:583-633, default options): anchors, pre-chain,
    each query's kept anchors sorted, chain, all on the device (nsh_index_chain; the queries'
    upload and the chains' download included). The reference's code (oracle/_ref) runs the same
    steps per query on one host thread over a sample; its chains must be identical"""
    import ctypes as C
    import numpy as np
    import nsffi
    n_u = int(sum(len(u) for u, _ in chains))
    out = {"queries": len(qs), "chains": n_u, "chain_anchors": int(sum(len(a) for _, a in chains)),
           "gpu_ms": round(gpu_s * 1e3, 1), "gpu_queries_per_s": round(len(qs) / gpu_s, 1)}
    if log:
        log("[bench] index -> chain: %d queries, %d chains, GPU %.1f ms" % (len(qs), n_u, gpu_s * 1e3))
    if not (cpu and nsffi.have_ref()):
        return out
    ref = nsffi.Reference()
    L = ref.lib
    h = L.or_ref_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, 30, 6, 1, 8)
    buf = np.zeros(1 << 23, np.uint64)
    u1, a1, u2, a2 = (np.zeros(1 << 23, np.uint64) for _ in range(4))
    same, done, t_ref = 0, 0, 0.0
    try:
        t_end = time.perf_counter() + 12.0
        for k, q in enumerate(qs):
            t0 = time.perf_counter()
            n = L.or_ref_anchors(h, q, len(q), 20000, len(buf), buf.ctypes.data)
            nu1 = L.or_ref_chain(256, 256, 256, 25, 1000000, 2, 0, 0.75, 1, 6, 8, n, buf.ctypes.data, len(u1),
                                 u1.ctypes.data, len(a1), a1.ctypes.data)
            m1 = int((u1[:nu1] & np.uint64(0xffffffff)).sum())
            a1[:m1].sort()
            nu2 = L.or_ref_chain(200000, 1000, 200000, 25, 1000000, 3, 0, 0.75, 1, 6, 8, m1, a1.ctypes.data,
                                 len(u2), u2.ctypes.data, len(a2), a2.ctypes.data)
            t_ref += time.perf_counter() - t0
            m2 = int((u2[:nu2] & np.uint64(0xffffffff)).sum())
            gu, ga = chains[k]
            same += np.array_equal(gu, u2[:nu2]) and np.array_equal(ga, a2[:m2])
            done += 1
            if time.perf_counter() > t_end:
                break
    finally:
        L.or_ref_index_free(h)
    out["cpu_baseline"] = {"value": round(t_ref * 1e3 / done * len(qs), 1),
                           "unit": "ms per %d queries (reference code, one host thread, scaled)" % len(qs),
                           "cores": 1, "kind": "reference",
                           "sample": "the first %d queries, %.1f s of reference work" % (done, t_ref)}
    out["identical_to_reference"] = "%d of %d sampled queries" % (same, done)
    out["gpu_vs_cpu_thread"] = round(out["cpu_baseline"]["value"] / out["gpu_ms"], 1)
    if log:
        log("[bench] index -> chain: reference %.1f ms (scaled), identical %s" % (out["cpu_baseline"]["value"],
                                                                                 out["identical_to_reference"]))
    return out


def bench_chain(anchors, cpu=True, log=None):
    """mp_map_coarse's two chaining passes (map.c:611-633, default options, spliced) over every
    query's anchors: the pre-chain (w = 256, min_cnt 2), each query's kept anchors sorted, then the
    main chain (max_intron 200000, max_gap 1000, bw 200000, max_skip 25, max_iter 1e6, min_cnt 3,
    coef 0.75). GPU times include the anchors' upload and the chains' download; the reference's
    mp_chain (oracle/_ref) runs the same two passes on one host thread over a sample of queries"""
    import numpy as np
    import nsffi
    from protlocus_amd import capi
    L = capi.lib()
    pre = (256, 256, 256, 25, 1000000, 2, 0, 0.75, 1, 6, 8)
    main = (200000, 1000, 200000, 25, 1000000, 3, 0, 0.75, 1, 6, 8)

    def run(lists, params):
        lens = np.array([len(a) for a in lists], np.int64)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        a = np.ascontiguousarray(np.concatenate(lists), dtype=np.uint64)
        n_u = np.zeros(len(lists), np.int32)
        u, ao = np.empty(max(1, len(a)), np.uint64), np.empty(max(1, len(a)), np.uint64)
        t0 = time.perf_counter()
        r = L.nsh_chain_batch(len(lists), off.ctypes.data, a.ctypes.data, *params, n_u.ctypes.data, u.ctypes.data,
                              ao.ctypes.data)
        dt = time.perf_counter() - t0
        assert r >= 0, capi.error()
        return off, n_u, u, ao, dt

    run(anchors, pre)   # warm-up: the library keeps its chaining workspace between calls
    off, n_u, u, ao, pre_s = run(anchors, pre)
    kept = []
    for q in range(len(anchors)):
        uq = u[off[q]:off[q] + n_u[q]]
        m = int((uq & np.uint64(0xffffffff)).sum())
        kept.append(np.sort(ao[off[q]:off[q] + m]))
    off2, n_u2, u2, ao2, main_s = run(kept, main)
    out = {"queries": len(anchors), "anchors": int(sum(len(a) for a in anchors)),
           "prechain_kept_anchors": int(sum(len(a) for a in kept)), "chains": int(n_u2.sum()),
           "gpu_prechain_ms": round(pre_s * 1e3, 1), "gpu_chain_ms": round(main_s * 1e3, 1)}
    if log:
        log("[bench] chain: %d -> %d anchors, %d chains; GPU pre-chain %.1f ms, chain %.1f ms" % (
            out["anchors"], out["prechain_kept_anchors"], out["chains"], pre_s * 1e3, main_s * 1e3))
    if cpu and nsffi.have_ref():
        L = nsffi.Reference().lib
        u1, a1, ur, ar = (np.zeros(1 << 23, np.uint64) for _ in range(4))
        same, done, cpu_s = 0, 0, 0.0
        t_end = time.perf_counter() + 10.0
        for q in range(len(anchors)):
            aq = np.ascontiguousarray(anchors[q])
            t0 = time.perf_counter()
            nu1 = L.or_ref_chain(*pre, len(aq), aq.ctypes.data, len(u1), u1.ctypes.data, len(a1), a1.ctypes.data)
            m1 = int((u1[:nu1] & np.uint64(0xffffffff)).sum())
            a1[:m1].sort()
            nu2 = L.or_ref_chain(*main, m1, a1.ctypes.data, len(ur), ur.ctypes.data, len(ar), ar.ctypes.data)
            cpu_s += time.perf_counter() - t0
            gu = u2[off2[q]:off2[q] + n_u2[q]]
            ga = ao2[off2[q]:off2[q] + int((gu & np.uint64(0xffffffff)).sum())]
            m2 = int((ur[:nu2] & np.uint64(0xffffffff)).sum())
            same += np.array_equal(gu, ur[:nu2]) and np.array_equal(ga, ar[:m2])
            done += 1
            if time.perf_counter() > t_end:
                break
        out["cpu_baseline"] = {"value": round(cpu_s * 1e3 / done * len(anchors), 1),
                               "unit": "ms per %d queries, both passes (reference code, one host thread, scaled)" % len(
                                   anchors), "cores": 1, "kind": "reference",
                               "sample": "the first %d queries, %.1f s of reference work" % (done, cpu_s)}
        out["identical_to_reference"] = "%d of %d sampled queries" % (same, done)
        if log:
            log("[bench] chain: reference %s, identical %s" % (out["cpu_baseline"]["value"],
                                                               out["identical_to_reference"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="pick each column group's persistent wave count on this batch first (untimed)")
    ap.add_argument("--no-extras", action="store_true", help="headline only (no prescreen / hmm sections)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only-profile-search", action="store_true",
                    help="run only the profile prescreen search section and print its JSON object")
    ap.add_argument("--only-gapless-prefilter", action="store_true",
                    help="run only the gapless prefilter section and print its JSON object")
    ap.add_argument("--only-index", action="store_true", help="run only the k-mer index section and print its JSON")
    ap.add_argument("--only-map", action="store_true",
                    help="run only the genome-scale map path section and print its JSON")
    ap.add_argument("--map-batch", type=int, default=20000, help="proteins per nsh_map_batch call")
    ap.add_argument("--only-map-prescreen", action="store_true",
                    help="run only the learned prescreen on the map path's candidates and print its JSON")
    ap.add_argument("--only-prescreen", type=int, default=0, metavar="PROTEINS",
                    help="run only the learned prescreen section with PROTEINS evaluation proteins (two candidates "
                         "each; 524288 gives the workload's ~1M candidates) and print its JSON")
    ap.add_argument("--only-dropin", action="store_true",
                    help="run only the drop-in section (reference calling patterns from C threads) and print its JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times (a copy kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    ap.add_argument("--scalar-read-calibration", type=int, default=0, metavar="MIB",
                    help="only read MIB MiB through the scalar data cache three times (nsh_diag_scalar_read: the "
                         "path of the extension DP's row records) and exit")
    args = ap.parse_args()
    if args.scalar_read_calibration:
        hip = C.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        hip.hipFree.argtypes = [C.c_void_p]
        from protlocus_amd import capi
        lib = capi.lib()
        lib.nsh_diag_scalar_read.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        n = args.scalar_read_calibration << 20
        src, out = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), n) or hip.hipMalloc(C.byref(out), n // 65536 * 4 + 4) or hip.hipMemset(src, 1, n):
            raise SystemExit("hipMalloc/hipMemset failed")
        hip.hipDeviceSynchronize()
        for _ in range(3):
            if not lib.nsh_diag_scalar_read(src, n, out):
                raise SystemExit("nsh_diag_scalar_read failed")
        hip.hipFree(src), hip.hipFree(out)
        print(json.dumps({"scalar_read_calibration_bytes": n, "launches": 3}), flush=True)
        return
    if args.copy_calibration:
        # HIP runtime device-to-device copies (the runtime's copy kernel): n bytes read, n written
        hip = C.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        hip.hipFree.argtypes = [C.c_void_p]
        n = args.copy_calibration << 20
        src, dst = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), n) or hip.hipMalloc(C.byref(dst), n) or hip.hipMemset(src, 1, n):
            raise SystemExit("hipMalloc/hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, n, 3):   # hipMemcpyDeviceToDevice
                raise SystemExit("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src), hip.hipFree(dst)
        print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; the DP problems are "
              "independent, so N GPUs would shard pairs by cost (longest-processing-time bins) - run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    if args.only_profile_search:
        print(json.dumps(bench_profile_search(cpu=not args.no_cpu, log=log)), flush=True)
        return
    if args.only_gapless_prefilter:
        print(json.dumps(bench_gapless_prefilter(cpu=not args.no_cpu, log=log)), flush=True)
        return
    if args.only_index:
        print(json.dumps(bench_index(cpu=not args.no_cpu, log=log)), flush=True)
        return
    if args.only_map:
        print(json.dumps(bench_map(cpu=not args.no_cpu, log=log, batch=args.map_batch)), flush=True)
        return
    if args.only_map_prescreen:
        from protlocus_amd import prescreen
        print(json.dumps(prescreen.map_evaluate(log=log)), flush=True)
        return
    if args.only_prescreen:
        from protlocus_amd import prescreen
        ps = prescreen.evaluate(n_train=32768, n_val=8192, n_eval=args.only_prescreen, epochs=20, seed=11, log=log)
        ps["workload"] = ("%d proteins x (true locus + decoy window), protein 50-1000 aa; trained on 32,768 other "
                          "proteins' candidates; DP = plain global spliced DP over the window" % args.only_prescreen)
        print(json.dumps(ps), flush=True)
        return

    import nsffi
    from protlocus_amd import synth
    t0 = time.time()
    pairs = synth.extension_batch(args.pairs, seed=args.seed)
    if args.only_dropin:
        prod = nsffi.Product()
        if not prod.lib.ns_gpu_init(0):
            raise SystemExit("GPU not usable: " + prod.error())
        kw = dict(flag=nsffi.NS_F_EXT_RIGHT)
        packed = synth.pack_batch(pairs)
        ref_res = None if args.no_cpu else cpu_baseline(pairs, kw, seconds=args.cpu_seconds, packed=packed)[1]
        print(json.dumps(bench_dropin(prod, pairs, packed, kw, ref_res, log, cpu=not args.no_cpu)), flush=True)
        return
    log("[bench] generated %d pairs in %.1f s" % (len(pairs), time.time() - t0))
    prod = nsffi.Product()
    if not prod.lib.ns_gpu_init(0):
        raise SystemExit("GPU not usable: " + prod.error())
    kw = dict(flag=nsffi.NS_F_EXT_RIGHT)
    # a 64-pair plan first: the process's first plan creates the HIP streams and hardware queues
    # that later plans reuse, a one-time cost kept out of the timed build below
    w = synth.pack_batch(pairs[:64])
    prod.ext_plan(*w, prod.opt(**kw)).close()
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    t_up = time.perf_counter()
    plan = prod.ext_plan(ns, ns_off, nl, as_, as_off, al, prod.opt(**kw))   # host buffers -> HBM (PCIe), once
    plan_build_s = time.perf_counter() - t_up
    waves, tune_s = None, 0.0
    if args.autotune:   # persistent waves per column-count group, picked on this batch (untimed)
        t0 = time.perf_counter()
        waves = plan.autotune(reps=3)
        tune_s = time.perf_counter() - t0
        log("[bench] autotuned waves per CU %s, pairs up to %d rows, in %.1f s" % (waves, plan.pair_rows(), tune_s))
    for _ in range(args.warmup):
        plan.run()
    log("[bench] warmup done")
    prod.lib.nsh_device_synchronize()
    dp_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(sync=False)
        plan.prod.lib.nsh_ext_plan_sync(plan.p)
        dp_ms.append(plan.ms(1))
    prod.lib.nsh_device_synchronize()
    wall = time.perf_counter() - t0
    gpu_res = plan.results()
    cells = plan.cells(1)          # evaluated cells of the last run
    win = plan.cells(0)
    rows_done = plan.cells(2)      # evaluated rows (the row records the DP reads)
    # the same step replayed as one captured hipGraph (preparation, both group kernels): its wall
    # clock per step and whether every replay's results equal the stream-launched run's
    plan.run(graph=True)           # capture (untimed)
    prod.lib.nsh_device_synchronize()
    g_same, t0 = True, time.perf_counter()
    for _ in range(args.steps):
        plan.run(sync=False, graph=True)
        plan.prod.lib.nsh_ext_plan_sync(plan.p)
    prod.lib.nsh_device_synchronize()
    g_ms = (time.perf_counter() - t0) / args.steps * 1e3
    g_same = plan.results() == gpu_res
    pair_rows = plan.pair_rows()
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
                   "pairs": args.pairs, "seed": args.seed, "cell_definition": "(rows evaluated) x al",
                   "waves_per_cu": waves if waves else "compiled defaults",
                   "pair_rows": pair_rows,
                   "autotune_s": round(tune_s, 2)},
        "proteins_per_s": round(args.pairs / (ms * 1e-3), 1),
        "window_gcups": round(win / (ms * 1e-3) / 1e9, 2),
        "dp_kernel_ms": round(dp_avg, 3),
        # one step's work when the inputs start in host memory: the plan build (packing, PCIe upload)
        # added to a step; never the headline
        "graph_replay": {"ms_per_step": round(g_ms, 3), "gcups": round(cells / (g_ms * 1e-3) / 1e9, 2),
                         "identical_to_stream_run": g_same},
        "pcie_inclusive": {"plan_build_ms": round(plan_build_s * 1e3, 2),
                           "gcups": round(cells / ((ms + plan_build_s * 1e3) * 1e-3) / 1e9, 2)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                     "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": round(ext_dp_traffic(rows_done, args.pairs), -5)
                     if (args.pairs, args.seed) == (65536, 1) else None,
                     "ops_per_cell": OPS_PER_CELL},
    }
    log("[bench] gpu: %.2f GCUPS, %.1f ms/step, dp %.1f ms; %.0f rows evaluated" % (gcups, ms, dp_avg, rows_done))
    ref_res = None
    if not args.no_cpu:
        cb, ref_res = cpu_baseline(pairs, kw, seconds=args.cpu_seconds, packed=(ns, ns_off, nl, as_, as_off, al))
        out["cpu_baseline"] = cb
        if ref_res is not None:   # every pair's (nt_len, aa_len, score) against the reference's
            same = sum(1 for k, g in enumerate(gpu_res) if tuple(int(v) for v in ref_res[k]) == g)
            out["identical_to_reference"] = "%d of %d" % (same, len(gpu_res))
            log("[bench] identical to the reference: %d of %d pairs" % (same, len(gpu_res)))
            if same != len(gpu_res):
                bad = [k for k, g in enumerate(gpu_res) if tuple(int(v) for v in ref_res[k]) != g][:5]
                raise SystemExit("headline results differ from the reference's on %d pairs, e.g. %s" %
                                 (len(gpu_res) - same, [(k, gpu_res[k], tuple(ref_res[k])) for k in bad]))
        out["gpu_window_vs_cpu"] = round(out["window_gcups"] / cb["value"], 1) if cb["value"] else None
        log("[bench] cpu baseline: %.2f GCUPS (%s, %d threads)" % (cb["value"], cb["kind"], cb["cores"]))
    if not args.no_extras:
        out["drop_in"] = bench_dropin(prod, pairs, (ns, ns_off, nl, as_, as_off, al), kw,
                                      ref_res if not args.no_cpu else None, log, cpu=not args.no_cpu)
        from protlocus_amd import prescreen
        del pairs
        t0 = time.time()
        # 131,072 candidates: toward the ~1M of the prescreen workload while the run stays within a
        # minute (the pipeline gain grows with the batch: the DP's tail weighs less)
        ps = prescreen.evaluate(n_train=32768, n_val=8192, n_eval=65536, epochs=20, seed=11, log=log)
        ps["workload"] = ("65,536 proteins x (true locus + decoy window), protein 50-1000 aa; trained on 32,768 "
                          "other proteins' candidates; DP = plain global spliced DP over the window")
        n = ps["candidates"]
        if ps["mlp_infer_ms"] > 0:
            tf = n * MLP_FLOP_PER_SAMPLE / (ps["mlp_infer_ms"] * 1e-3) / 1e12
            ps["mfma"] = {"achieved": round(tf, 2), "peak": MFMA_BF16_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(tf / MFMA_BF16_TFLOPS, 5), "flop_per_candidate": MLP_FLOP_PER_SAMPLE}
        out["prescreen"] = ps
        log("[bench] prescreen: recall %.4f, kept %.3f, speedup %.2fx (%.1f s)" % (ps["recall"], ps["kept_frac"],
                                                                                  ps["speedup"], time.time() - t0))
        out["hmm"] = bench_hmm(log=log)
        log("[bench] hmm: %.1f GCUPS" % out["hmm"]["gcups"])
        out["profile_search"] = bench_profile_search(cpu=not args.no_cpu, log=log)
        out["gapless_prefilter"] = bench_gapless_prefilter(cpu=not args.no_cpu, log=log)
        out["index"] = bench_index(cpu=not args.no_cpu, log=log)
        out["map"] = bench_map(cpu=not args.no_cpu, log=log, batch=args.map_batch)
        out["map_prescreen"] = prescreen.map_evaluate(log=log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
