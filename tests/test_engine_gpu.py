"""GPU parity of the batching engines and of long proteins, against the oracle (bit-exact).

  * long proteins: extensions with padded widths over the one-wave kernel's 2048 columns (the
    work-group kernel in extension mode, up to 8192), left and right, matrix and PSSM scores, one by
    one and mixed into a batch with short ones; global CIGAR DP above 4096 columns (up to 8192),
    int16 and int32. The reference's CPU path has no length cap (nasw-sse.c:763-772, 779-973);
  * the asynchronous ns_gpu_submit / ns_gpu_wait: submissions from several threads, waited in any
    order, are batched (fewer launches than requests) and each result equals the oracle's;
  * concurrent synchronous calls (ns_global_gs16b / gs32b with NS_F_CIGAR from several threads, as
    kt_for workers make them) share batches and give every thread its own result;
  * the trace budget: a batch whose traces exceed it is cut into several parts with equal results.
"""
import concurrent.futures as cf
import ctypes as C
import os

import numpy as np
import pytest

import nsffi
from cases import planted_case, random_opt_kw, random_ss
from nsffi import NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT

pytestmark = pytest.mark.gpu


def _ora_ext(oracle, product, ns, a, kw, ss=None):
    w = product.lib.ns_ext_stripe_width(C.byref(product.opt(**kw)), len(a))
    return oracle.ext(ns, a, oracle.opt(**kw), ss, (len(a) + w - 1) // w * w)[0]


def test_long_extensions(oracle, product):
    rng = np.random.default_rng(2049)
    for al in (2033, 2049, 2100, 3000, 4100, 5000, 8000):
        ns, a = planted_case(rng, al)
        kw = dict(flag=NS_F_EXT_RIGHT if al % 2 else NS_F_EXT_LEFT, **random_opt_kw(rng))
        if al in (3000, 8000):
            kw["pssm"] = rng.integers(-6, 9, size=(al, 22)).astype(np.int8)
        ss = random_ss(rng, len(ns)) if al == 4100 else None
        print("long extension", al, kw["flag"], "pssm" in kw, flush=True)
        got = product.run("ns_global_gs16b", ns, a, product.opt(**kw), ss)[:3]
        assert got == _ora_ext(oracle, product, ns, a, kw, ss), (al, kw["flag"])
    # plain score-only global DP (mode 1) of a long protein
    ns, a = planted_case(rng, 2600)
    exp = oracle.global_dp(ns, a, oracle.opt(flag=0), None, mode=16)
    assert product.run("ns_global_gs16b", ns, a, product.opt(flag=0)) == exp


def test_long_and_short_requests_in_one_batch(oracle, product):
    """requests over 2048 columns share the flush (and the option group) with short ones; every
    result is correct (none falls back to INT32_MIN)"""
    L = product.lib
    rng = np.random.default_rng(7)
    b = L.ns_ext_batch_init(None)
    reqs, keep = [], []
    for k in range(24):
        al = int(rng.integers(2049, 3500)) if k % 4 == 0 else int(rng.integers(20, 900))
        ns, a = planted_case(rng, al)
        kw = dict(flag=NS_F_EXT_LEFT if k & 1 else NS_F_EXT_RIGHT)
        o = product.opt(**kw)
        ab = a.encode()
        keep.append((o, ab))
        reqs.append((L.ns_ext_batch_add(b, ns, len(ns), ab, len(a), C.byref(o), None), ns, a, kw))
    assert L.ns_ext_batch_flush(b) == len(reqs)
    for rid, ns, a, kw in reqs:
        r = L.ns_ext_batch_get(b, rid).contents
        assert (r.nt_len, r.aa_len, r.score) == _ora_ext(oracle, product, ns, a, kw), (len(a), kw["flag"])
    L.ns_ext_batch_destroy(b)
    # a plan of long and short problems, also through a subset run
    from protlocus_amd import synth
    pairs = [planted_case(rng, int(al)) for al in (2100, 60, 3000, 300, 2500, 900)]
    dirs = np.array([0, 1, 1, 0, 1, 0], dtype=np.uint8)
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch([(d, p.encode()) for d, p in pairs])
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, product.opt(flag=NS_F_EXT_RIGHT), dirs=dirs)
    try:
        plan.run()
        got = plan.results()
        keep_mask = np.array([1, 0, 1, 1, 0, 1], dtype=np.uint8)
        assert product.lib.nsh_ext_plan_run_subset(plan.p, keep_mask.tobytes()), product.error()
        product.lib.nsh_ext_plan_sync(plan.p)
        sub = plan.results()
    finally:
        plan.close()
    for k, (d, p) in enumerate(pairs):
        exp = _ora_ext(oracle, product, d, p, dict(flag=NS_F_EXT_LEFT if dirs[k] else NS_F_EXT_RIGHT))
        assert got[k] == exp, k
        assert sub[k] == (exp if keep_mask[k] else (0, 0, -2 ** 31)), k


@pytest.mark.parametrize("is32", [0, 1])
def test_long_global_cigar(oracle, product, is32):
    rng = np.random.default_rng(4097 + is32)
    for al in (4097, 5000, 8192):
        ns, a = planted_case(rng, al, flank=10)
        kw = dict(flag=NS_F_CIGAR, **random_opt_kw(rng))
        print("long global", al, is32, flush=True)
        exp = oracle.global_dp(ns, a, oracle.opt(**kw), None, mode=32 if is32 else 16)
        got = product.run("ns_global_gs32b" if is32 else "ns_global_gs16b", ns, a, product.opt(**kw))
        assert got == exp, (al, is32, got[:3], exp[:3])


def test_submit_from_threads_is_batched(oracle, product):
    L = product.lib
    rng = np.random.default_rng(31)
    cases = []
    for _ in range(96):
        ns, a = planted_case(rng, int(rng.integers(3, 700)))
        cases.append((ns, a))
    launches0 = L.ns_gpu_launch_count()

    def submit(k):
        ns, a = cases[k]
        o = product.opt(flag=NS_F_CIGAR)
        q = L.ns_gpu_submit(None, ns, len(ns), a.encode(), len(a), C.byref(o), None)
        assert q, product.error()
        return q
    with cf.ThreadPoolExecutor(8) as ex:
        qs = list(ex.map(submit, range(len(cases))))
    order = list(range(len(cases)))
    rng.shuffle(order)
    got = {}
    for k in order:
        r = nsffi.NsRst()
        assert L.ns_gpu_wait(qs[k], C.byref(r))
        got[k] = (r.nt_len, r.aa_len, r.score, nsffi.cigar_str(r))
        if r.cigar:
            product.free(C.cast(r.cigar, C.c_void_p))
    launches = L.ns_gpu_launch_count() - launches0
    assert launches < len(cases), launches          # batched: far fewer launches than requests
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(lambda c: oracle.global_dp(c[0], c[1], oracle.opt(flag=NS_F_CIGAR), mode=16), cases))
    for k in range(len(cases)):
        assert got[k] == exp[k], k


def test_concurrent_synchronous_calls(oracle, product):
    rng = np.random.default_rng(77)
    cases = []
    for k in range(64):
        ns, a = planted_case(rng, int(rng.integers(1, 600)))
        cases.append((ns, a, dict(flag=NS_F_CIGAR, **random_opt_kw(rng)), random_ss(rng, len(ns)), k % 3 == 0))

    def run(c):
        ns, a, kw, ss, is32 = c
        return product.run("ns_global_gs32b" if is32 else "ns_global_gs16b", ns, a, product.opt(**kw), ss)
    with cf.ThreadPoolExecutor(16) as ex:
        got = list(ex.map(run, cases))
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(lambda c: oracle.global_dp(c[0], c[1], oracle.opt(**c[2]), c[3], mode=32 if c[4] else 16),
                          cases))
    bad = [k for k in range(len(cases)) if got[k] != exp[k]]
    assert not bad, [(k, len(cases[k][1]), got[k][:3], exp[k][:3]) for k in bad[:4]]


def test_trace_budget_cuts_a_batch(oracle, product, monkeypatch):
    """with a 1-MB trace budget, 24 submissions of about 0.3-1.4 MB of trace each run in several
    parts of one batch; results are unchanged"""
    L = product.lib
    rng = np.random.default_rng(12)
    cases = [planted_case(rng, int(rng.integers(200, 420)), flank=200) for _ in range(24)]
    monkeypatch.setenv("NSH_TRACE_BUDGET_MB", "1")
    launches0 = L.ns_gpu_launch_count()
    qs = []
    for ns, a in cases:
        o = product.opt(flag=NS_F_CIGAR)
        qs.append(L.ns_gpu_submit(None, ns, len(ns), a.encode(), len(a), C.byref(o), None))
    got = []
    for q in qs:
        r = nsffi.NsRst()
        assert L.ns_gpu_wait(q, C.byref(r)), product.error()
        got.append((r.nt_len, r.aa_len, r.score, nsffi.cigar_str(r)))
        if r.cigar:
            product.free(C.cast(r.cigar, C.c_void_p))
    assert L.ns_gpu_launch_count() - launches0 >= 8    # at least two parts of prep, DP and backtrack
    for (ns, a), g in zip(cases, got):
        assert g == oracle.global_dp(ns, a, oracle.opt(flag=NS_F_CIGAR), mode=16)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_global_batch_entry(oracle, product, mode):
    """nsh_global_batch: one engine batch of 300 problems (int16, int32, or per problem as
    ns_global_gs_auto); scores and CIGARs in input order, bit-exact"""
    from protlocus_amd import synth
    L = product.lib
    L.nsh_global_batch.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p,
                                   C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.POINTER(C.c_uint32))]
    rng = np.random.default_rng(300 + mode)
    pairs = [planted_case(rng, int(rng.integers(1, 900)) if k % 50 else 2800) for k in range(300)]
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch([(d, p.encode()) for d, p in pairs])
    ns_off, as_off = (np.ascontiguousarray(x, dtype=np.int64) for x in (ns_off, as_off))
    nl, al = (np.ascontiguousarray(x, dtype=np.int32) for x in (nl, al))
    score = np.zeros(len(pairs), dtype=np.int32)
    ncig = np.zeros(len(pairs), dtype=np.int32)
    cig = C.POINTER(C.c_uint32)()
    o = product.opt(flag=NS_F_CIGAR)
    assert L.nsh_global_batch(len(pairs), ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data,
                              al.ctypes.data, None, C.byref(o), mode, score.ctypes.data, ncig.ctypes.data,
                              C.byref(cig)), product.error()
    off = 0
    ops = "MIDNSHP=XBFGUVE"
    got = []
    for k in range(len(pairs)):
        s = "".join("%d%s" % (cig[off + t] >> 4, ops[cig[off + t] & 0xf]) for t in range(ncig[k]))
        got.append((int(score[k]), s))
        off += int(ncig[k])
    product.free(C.cast(cig, C.c_void_p))
    m = {0: 16, 1: 32, 2: "auto"}[mode]
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(lambda pr: oracle.global_dp(pr[0], pr[1], oracle.opt(flag=NS_F_CIGAR), mode=m), pairs))
    bad = [k for k in range(len(pairs)) if got[k] != (exp[k][2], exp[k][3])]
    assert not bad, [(k, len(pairs[k][1]), got[k][0], exp[k][2]) for k in bad[:4]]
