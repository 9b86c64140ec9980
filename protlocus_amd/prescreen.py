"""Learned chain-candidate prescreen: candidates, exact-DP labels, training on the GPU, and the
prescreened DP pipeline.

A candidate is a (protein, DNA window) pair as chaining hands it to the spliced DP. Its 64
features come from the GPU feature kernel (k-mer anchors between the protein and the window's
three frames, diagonal statistics, splice-site and composition counts). Labels come from the
exact DP itself: the plain score-only global spliced DP over the window (ns_global_gs16b
semantics, every row evaluated) accepts a candidate when score >= 2*al - 150. The network
(64 -> 256 -> 256 -> 1, BF16 MFMA, FP32 Adam) is trained with the library's own forward and
backward kernels; a logit threshold is calibrated on held-out candidates for >= 99.8% recall
of DP-accepted candidates. At inference the network runs ahead of the batched DP on the
plan's resident inputs; rejected candidates are skipped except for a 2% random audit that
still goes through the full DP, which stays the arbiter of every accepted alignment.
"""
import time

import numpy as np

from . import capi, synth

ACCEPT_SLOPE, ACCEPT_OFFSET = 2, -150
LABEL_KW = dict(flag=0, xdrop=1 << 30)       # plain global DP, no x-drop cut: every row evaluated
AUDIT_FRAC = 0.02
TARGET_RECALL = 0.998                        # calibration target (the reported bar is 0.995)


def accept(scores, al):
    return scores >= ACCEPT_SLOPE * np.asarray(al, dtype=np.int64) + ACCEPT_OFFSET


def candidates(n_prot, seed, min_len=50, max_len=1000, log=None):
    """two candidates per protein: its true locus (10-45% substitutions, 0-3 introns, 0-2
    frameshifts, 200-nt flanks) and a decoy window of the same length: 70% random genomic DNA
    (GC 0.45), 30% a distant homolog (55-80% substitutions) that scores like random k-mer hits"""
    rng = np.random.default_rng(seed)
    prots, dnas, kind = [], [], []
    for k in range(n_prot):
        if log and k and k % 65536 == 0:
            log("[prescreen] %d of %d proteins' candidates made" % (k, n_prot))
        L = int(round(np.exp(rng.uniform(np.log(min_len), np.log(max_len)))))
        prot = synth.rand_protein(rng, L)
        ni, nfs = int(rng.integers(0, 4)), int(rng.integers(0, 3))
        ilen = lambda: min(5000, max(8, rng.lognormal(np.log(300), 0.8)))
        true = synth.gene_of(rng, prot, rng.uniform(0.10, 0.45), ni, ilen, nfs, 200)
        if rng.random() < 0.7:
            decoy = synth.rand_dna(rng, len(true), gc=0.45)
        else:
            decoy = synth.gene_of(rng, prot, rng.uniform(0.55, 0.80), ni, ilen, nfs, 200)
        prots += [prot, prot]
        dnas += [true, decoy]
        kind += [1, 0]
    return prots, dnas, np.array(kind, dtype=np.int8)


class Batch:
    """candidates resident on the device with their DP labels and features"""

    def __init__(self, prots, dnas):
        self.prots, self.dnas = prots, dnas
        self.al = np.array([len(p) for p in prots], dtype=np.int64)
        self.nl = np.array([len(d) for d in dnas], dtype=np.int64)
        self.plan = capi.Plan(dnas, prots, capi.opt(**LABEL_KW))
        self.plan.run()
        self.scores = self.plan.scores()
        self.labels = accept(self.scores, self.al)
        self.cells = self.nl * self.al

    def close(self):
        self.plan.close()


def init_params(rng):
    """He-initialised FP32 parameter vector in the library's layout"""
    W1 = rng.normal(0, np.sqrt(2 / 64), size=(64, 256))
    W2 = rng.normal(0, np.sqrt(2 / 256), size=(256, 256))
    w3 = rng.normal(0, np.sqrt(1 / 256), size=256)
    return np.concatenate([W1.ravel(), np.zeros(256), W2.ravel(), np.zeros(256), w3, [0.0]]).astype(np.float32)


def train(X, y, epochs=12, batch=1024, lr=2e-3, seed=0, pos_weight=1.0, log=None):
    """train on the GPU; returns (FP32 parameters, feature means, feature sds). pos_weight > 1 draws accepted
    candidates that much more often (the recall target makes the low tail of their logits matter)"""
    rng = np.random.default_rng(seed)
    mu = X.mean(axis=0).astype(np.float32)
    sd = (X.std(axis=0) + 1e-3).astype(np.float32)
    Xs = np.ascontiguousarray(X, dtype=np.float32)      # standardised on the device
    net = capi.MLP(init_params(rng), mu, sd)
    n = len(y) // batch * batch
    w = np.where(y > 0, pos_weight, 1.0)
    w = w / w.sum()
    try:
        for ep in range(epochs):
            idx = rng.choice(len(y), size=n, replace=True, p=w) if pos_weight != 1.0 else rng.permutation(len(y))[:n]
            tot = 0.0
            step_lr = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
            for s in range(0, n, batch):
                b = idx[s:s + batch]
                tot += net.train_step(Xs[b], y[b].astype(np.float32), step_lr)
            if log:
                log("epoch %d loss %.4f" % (ep, tot / (n // batch)))
        params = net.params()
    finally:
        net.close()
    return params, mu, sd


def calibrate(logits, labels, target=TARGET_RECALL):
    """largest threshold keeping >= target of the positives"""
    pos = np.sort(logits[labels.astype(bool)])
    if len(pos) == 0:
        return float(logits.min())
    k = int(np.floor((1 - target) * len(pos)))
    return float(pos[k]) - 1e-4


def host_features(plan, n):
    """the plan's features (computed on the device) copied to the host, n x 64"""
    return plan.features_host()


def run_prescreened(plan, net, tau, n, rng, stages=None):
    """features -> logits -> keep (plus the audit) -> DP on the kept candidates; returns keep mask.
    stages (a dict) receives the wall time of each stage in ms."""
    t0 = time.perf_counter()
    ptr = plan.features_device()
    t1 = time.perf_counter()
    logit = net.infer(device_ptr=ptr, n=n)
    t2 = time.perf_counter()
    keep = logit >= tau
    rejected = np.nonzero(~keep)[0]
    audit = rng.choice(rejected, size=int(round(AUDIT_FRAC * len(rejected))), replace=False) if len(rejected) else []
    keep[audit] = True
    t3 = time.perf_counter()
    plan.run(keep=keep.astype(np.uint8))
    t4 = time.perf_counter()
    if stages is not None:
        for k, a, b in (("features", t0, t1), ("mlp", t1, t2), ("select", t2, t3), ("dp", t3, t4)):
            stages[k] = stages.get(k, 0.0) + (b - a) * 1e3
    return keep, logit


def evaluate(n_train=16384, n_val=4096, n_eval=16384, seed=11, epochs=12, pos_weight=1.0, lr=2e-3, log=print):
    """end-to-end: train, calibrate, then time the plain and the prescreened pipelines on a
    fresh candidate set; returns a dict of measurements"""
    t0 = time.time()
    tr = Batch(*candidates(n_train, seed)[:2])
    Xtr = host_features(tr.plan, len(tr.prots))
    ytr = tr.labels.astype(np.float32)
    tr.close()
    va = Batch(*candidates(n_val, seed + 1)[:2])
    Xva = host_features(va.plan, len(va.prots))
    if log:
        log("[prescreen] data %.1f s: train %d (%.1f%% accepted), val %d" % (time.time() - t0, len(ytr),
                                                                           100 * ytr.mean(), len(va.prots)))
    params, mu, sd = train(Xtr, ytr, epochs=epochs, seed=seed, pos_weight=pos_weight, lr=lr, log=log)
    net = capi.MLP(params, mu, sd)
    lv = net.infer(Xva)
    tau = calibrate(lv, va.labels)
    # hand-feature baseline: the best single feature threshold at the same calibration target
    best_f, best_keep = None, 1.0
    for f in range(capi.N_FEATURES):
        th = calibrate(Xva[:, f], va.labels)
        kf = (Xva[:, f] >= th).mean()
        if kf < best_keep:
            best_f, best_keep, best_th = f, kf, th
    va.close()
    ev = Batch(*candidates(n_eval, seed + 2, log=log)[:2])
    if log:
        log("[prescreen] evaluation set: %d candidates on the device" % len(ev.prots))
    n = len(ev.prots)
    rng = np.random.default_rng(seed + 3)
    # plain pipeline: the DP of every candidate (mean of 3 timed runs after a warm-up)
    ev.plan.run()
    t = time.perf_counter()
    for _ in range(3):
        ev.plan.run()
    plain_s = (time.perf_counter() - t) / 3
    accepted = accept(ev.plan.scores(), ev.al)    # refreshes the plan's per-problem results
    assert np.array_equal(accepted, ev.labels)
    cells_plain = ev.plan.cells(1)
    # prescreened pipeline
    run_prescreened(ev.plan, net, tau, n, rng)
    stages = {}
    t = time.perf_counter()
    for _ in range(3):
        keep, logit = run_prescreened(ev.plan, net, tau, n, rng, stages)
    pre_s = (time.perf_counter() - t) / 3
    got = ev.plan.scores()
    cells_pre = ev.plan.cells(1)
    # the same pipeline with a perfect classifier (keep = DP-accepted, plus the audit): the
    # ceiling any prescreen reaches on this candidate mix
    ideal = accepted.copy()
    rej = np.nonzero(~ideal)[0]
    ideal[rng.choice(rej, size=int(round(AUDIT_FRAC * len(rej))), replace=False)] = True
    t = time.perf_counter()
    for _ in range(3):
        ev.plan.features_device()
        ev.plan.run(keep=ideal.astype(np.uint8))
    ideal_s = (time.perf_counter() - t) / 3
    kept_pass = keep & accept(got, ev.al)
    recall = kept_pass[accepted].mean() if accepted.any() else 1.0
    Xev = host_features(ev.plan, n)
    base_keep = Xev[:, best_f] >= best_th
    out = {
        "candidates": n, "proteins": n // 2, "accepted_frac": round(float(accepted.mean()), 4),
        "tau": round(tau, 4), "kept_frac": round(float(keep.mean()), 4),
        "recall": round(float(recall), 5),
        "cells_saved": round(1 - cells_pre / cells_plain, 4),
        "plain_ms": round(plain_s * 1e3, 3), "prescreened_ms": round(pre_s * 1e3, 3),
        "speedup": round(plain_s / pre_s, 3),
        "perfect_classifier_speedup": round(plain_s / ideal_s, 3),
        "plain_proteins_per_s": round((n // 2) / plain_s, 1),
        "prescreened_proteins_per_s": round((n // 2) / pre_s, 1),
        "mlp_infer_ms": round(float(net.last_ms()), 4),
        "prescreened_stage_ms": {k: round(v / 3, 3) for k, v in stages.items()},
        "baseline_feature": int(best_f), "baseline_kept_frac": round(float(base_keep.mean()), 4),
        "baseline_recall": round(float(base_keep[accepted].mean()), 5),
        "train_candidates": int(len(ytr)),
    }
    net.close()
    ev.close()
    return out


# ---------------------------------------------------------------------------------------------
# the prescreen on the map path's own candidates
# ---------------------------------------------------------------------------------------------

def map_candidates(mapper, names, seqs):
    """every candidate the map path offers between coarse chaining and refinement, with its 64
    features, whether it became a record (the reference's own keep rule: the output step's filters,
    map.c:1143-1166), and its query index; one plain nsh_map_batch call"""
    import ctypes as C
    cap = 64 * len(seqs) + 1024
    while True:
        feat = np.zeros((cap, capi.N_FEATURES), np.float32)
        lab = np.zeros(cap, np.uint8)
        o = capi.MapOpt.default(cand_feat=feat.ctypes.data, cand_label=lab.ctypes.data, cand_cap=cap)
        recs, st = mapper.map(names, seqs, o)
        n = int(st[6])
        if n <= cap:
            break
        cap = n
    return feat[:n], lab[:n].astype(np.float32), recs, st


def map_evaluate(n_prot=20000, seed_train=22, seed_eval=21, epochs=30, lr=2e-3, log=print, target=TARGET_RECALL,
                 audit=AUDIT_FRAC):
    """train the network on a genome-scale workload's candidates, calibrate its cut on held-out
    queries, then map a second workload plain and prescreened: records, recall, time"""
    t0 = time.time()
    cn, cs, pn, ps, _ = synth.genome_with_genes(n_prot=n_prot, seed=seed_train)
    mp = capi.Mapper(cn, cs)
    try:
        X, y, _, st = map_candidates(mp, pn, ps)
    finally:
        mp.close()
    del cs
    n = len(y)
    if log:
        log("[map prescreen] training workload: %d candidates (%.1f%% became records) in %.1f s" %
            (n, 100 * y.mean(), time.time() - t0))
    cut = int(0.8 * n)   # candidates are in query order: the last fifth belongs to other queries
    tr, va = np.arange(cut), np.arange(cut, n)
    # a few training settings, the one that keeps the fewest held-out candidates at the target wins
    t1 = time.time()
    best = None
    for ep, pw in ((epochs, 1.0), (2 * epochs, 1.0), (2 * epochs, 2.0), (4 * epochs, 2.0)):
        params, mu, sd = train(X[tr], y[tr], epochs=ep, seed=seed_train, lr=lr, pos_weight=pw, log=None)
        cand = capi.MLP(params, mu, sd)
        lv = cand.infer(X[va])
        tau = calibrate(lv, y[va], target)
        kf = float((lv >= tau).mean())
        if log:
            log("[map prescreen] epochs %d, positive weight %.0f: keeps %.4f of held-out candidates" % (ep, pw, kf))
        if best is None or kf < best[0]:
            if best is not None:
                best[1].close()
            best = (kf, cand, tau, lv, ep, pw)
        else:
            cand.close()
    train_s = time.time() - t1
    _, net, tau, lv, best_ep, best_pw = best
    best_f, best_keep, best_th = None, 1.0, None
    for f in range(capi.N_FEATURES):
        for sign in (1, -1):
            th = calibrate(sign * X[va, f], y[va], target)
            kf = float((sign * X[va, f] >= th).mean())
            if kf < best_keep:
                best_f, best_keep, best_th = (f, sign), kf, th
    if log:
        log("[map prescreen] trained on %d candidates in %.1f s; tau %.3f keeps %.3f of held-out candidates "
            "(best single feature %s: %.3f)" % (len(tr), train_s, tau, float((lv >= tau).mean()), best_f, best_keep))
    # the evaluation workload: plain, then prescreened
    cn, cs, pn, ps, _ = synth.genome_with_genes(n_prot=n_prot, seed=seed_eval)
    mp = capi.Mapper(cn, cs)
    try:
        Xe, ye, _, _ = map_candidates(mp, pn, ps)      # untimed: features and labels of every candidate
        mp.map(pn[:32], ps[:32])
        t = time.perf_counter()
        plain, st_p = mp.map(pn, ps)
        plain_s = time.perf_counter() - t
        o = capi.MapOpt.default(prescreen=net.h, ps_tau=tau, ps_audit=audit, ps_seed=seed_eval)
        mp.map(pn[:32], ps[:32], o)
        t = time.perf_counter()
        pre, st_s = mp.map(pn, ps, o)
        pre_s = time.perf_counter() - t
        infer_ms = float(net.last_ms())
    finally:
        mp.close()
        net.close()
    a = set(plain.split(b"\n")) - {b""}
    b = set(pre.split(b"\n")) - {b""}
    recall = len(a & b) / max(1, len(a))
    sign = best_f[1]
    base_keep = float((sign * Xe[:, best_f[0]] >= best_th).mean())
    base_recall = float((sign * Xe[ye > 0, best_f[0]] >= best_th).mean()) if (ye > 0).any() else 1.0
    offered, dropped, audited, caught = int(st_s[6]), int(st_s[7]), int(st_s[14]), int(st_s[15])
    flop = 2 * (64 * 256 + 256 * 256 + 256)
    out = {
        "workload": "map path, %d proteins against 100 Mb (the map section's workload); trained on the "
                    "candidates of a second such workload (seed %d), cut calibrated on its last fifth" % (n_prot, seed_train),
        "candidates": offered, "became_records_frac": round(float(ye.mean()), 4),
        "kept_frac": round((offered - dropped) / max(1, offered), 4), "audited": audited,
        "audited_that_became_records": caught, "tau": round(tau, 4),
        "record_recall": round(recall, 5), "records_plain": len(a), "records_prescreened": len(b),
        "plain_s": round(plain_s, 3), "prescreened_s": round(pre_s, 3), "speedup": round(plain_s / pre_s, 3),
        "plain_proteins_per_s": round(n_prot / plain_s, 1), "prescreened_proteins_per_s": round(n_prot / pre_s, 1),
        "dp_candidates_plain": int(st_p[3]), "dp_candidates_prescreened": int(st_s[3]),
        "prescreen_ms": round((st_s[16] - st_s[9]) / 1e3, 1),
        "baseline_feature": "%d%s" % (best_f[0], "" if sign > 0 else " (low)"),
        "baseline_kept_frac": round(base_keep, 4), "baseline_candidate_recall": round(base_recall, 5),
        "train_candidates": int(len(tr)), "train_s": round(train_s, 2), "epochs": best_ep, "positive_weight": best_pw,
        "mlp_infer_ms": round(infer_ms, 4),
        "mfma": {"achieved": round(offered * flop / max(infer_ms, 1e-6) / 1e9, 3), "unit": "TFLOP/s",
                 "peak": 2500.0, "flop_per_candidate": flop},
    }
    out["mfma"]["frac"] = round(out["mfma"]["achieved"] / 2500.0, 5)

    def stages(st):
        prev, d = 0, {}
        for k, name in enumerate(capi.MAP_STAGES):
            d[name] = round((int(st[8 + k]) - prev) / 1e3, 1)
            prev = int(st[8 + k])
        return d
    out["stage_ms_plain"], out["stage_ms_prescreened"] = stages(st_p), stages(st_s)
    if log:
        log("[map prescreen] kept %.3f of %d candidates, record recall %.5f (%d -> %d records), %.2f s -> %.2f s "
            "(%.2fx); best single feature keeps %.3f" % (out["kept_frac"], offered, recall, len(a), len(b), plain_s,
                                                         pre_s, plain_s / pre_s, base_keep))
    return out
