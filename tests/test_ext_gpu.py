"""GPU parity of the batched extension DP against the oracle.

Mirrors the reference's extension tests: the closed forms of test/test_dp32.c:33-54 and the
mixed request set of test/test_gpu.c:292-386 (LCG inputs, AVX2/SSE stripe widths, with and
without a splice-score track, io 20/29, x-drop 30/100/1e6, end bonus 5/0, left and right,
several flushes in flight), plus seeded config-1-style batches. Bar: bit-exact
(nt_len, aa_len, score) on every request.
"""
import concurrent.futures as cf
import numpy as np
import pytest

import nsffi
from nsffi import NS_F_EXT_LEFT, NS_F_EXT_RIGHT, NS_S_GENERIC

pytestmark = pytest.mark.gpu


def lcg(seed):
    return (seed * 1103515245 + 12345) & 0xffffffff


def test_ext_closed_form(product):
    # TGG^n vs W^n: score = 11n + end_bonus, nt_len = 3n, aa_len = n (test/test_dp32.c:33-54)
    for flag in (NS_F_EXT_LEFT, NS_F_EXT_RIGHT):
        for n in (7, 19, 100, 700):
            o = product.opt(flag=flag)
            nt_len, aa_len, score, _ = product.run("ns_global_gs16b", b"TGG" * n, "W" * n, o)
            assert (nt_len, aa_len, score) == (3 * n, n, 11 * n + 5), (flag, n)


def ref_style_requests(seed=11, n_req=42, max_nt=30000, max_aa=700):
    """the request mix of test/test_gpu.c:300-345, restated"""
    aa_lens = [1, 3, 7, 15, 16, 17, 31, 33, 64, 100, 129, 257, 300, 511, 640]
    nt = bytearray(max_nt)
    aa = bytearray(max_aa)
    ss = bytearray(max_nt)
    for i in range(max_nt):
        seed = lcg(seed); nt[i] = b"ACGT"[(seed >> 16) & 3]
    for i in range(max_aa):
        seed = lcg(seed); aa[i] = b"ACDEFGHIKLMNPQRSTVWY"[(seed >> 16) % 20]
    for i in range(max_nt):
        seed = lcg(seed)
        ss[i] = ((((seed >> 20) % 40) + 44) << 1 | ((seed >> 27) & 1)) & 0xff if (seed >> 16) % 5 == 0 else 0xff
    codon = ["GCT", "TGT", "GAT", "GAA", "TTT", "GGT", "CAT", "ATT", "AAA", "TTA", "ATG", "AAT", "CCT", "CAA", "CGT",
             "TCT", "ACT", "GTT", "TGG", "TAT"]
    for k in range(400):
        c = codon["ACDEFGHIKLMNPQRSTVWY".index(chr(aa[k]))]
        p = 100 + k * 3 + (90 if k > 150 else 0)
        seed = lcg(seed)
        if p + 2 < max_nt and (seed >> 16) % 8:
            nt[p:p + 3] = c.encode()
    reqs = []
    for grp in range(3):
        for i in range(grp * n_req // 3, (grp + 1) * n_req // 3):
            seed = lcg(seed)
            kw = dict(sp_model=NS_S_GENERIC, flag=NS_F_EXT_LEFT if i & 1 else NS_F_EXT_RIGHT,
                      end_bonus=5 if grp == 0 else 0, xdrop=30 if grp == 1 else (1000000 if grp == 2 else 100))
            if i % 3 == 0:
                kw["io"] = 20
            al = aa_lens[i % len(aa_lens)]
            nl = 3 + (seed >> 8) % (max_nt - 10)
            if i % 6 == 0:
                nl = 3 + i // 6
            o_ns, o_as = (seed >> 4) % 7, (seed >> 12) % 5
            s = bytes(ss[o_ns:o_ns + nl]) if i % 3 == 1 else None
            reqs.append((grp, bytes(nt[o_ns:o_ns + nl]), bytes(aa[o_as:o_as + al]), kw, s))
    return reqs


def _oracle_ext(oracle, product, ns, a, kw, ss):
    o = oracle.opt(**kw)
    w = product.lib.ns_ext_stripe_width(nsffi.C.byref(product.opt(**kw)), len(a))
    npad = (len(a) + w - 1) // w * w
    return oracle.ext(ns, a, o, ss, npad)[0]


def test_ext_single_calls_match_oracle(oracle, product):
    for grp, ns, a, kw, ss in ref_style_requests():
        exp = _oracle_ext(oracle, product, ns, a, kw, ss)
        got = product.run("ns_global_gs16b", ns, a, product.opt(**kw), ss)[:3]
        assert got == exp, (grp, len(ns), len(a), kw, ss is not None)


def test_ext_batch_api_matches_oracle(oracle, product):
    L = product.lib
    reqs = ref_style_requests()
    b = L.ns_ext_batch_init(None)
    assert b, product.error()
    keep = []
    for grp in range(3):
        ids = []
        before = L.ns_ext_gpu_request_count()
        for g, ns, a, kw, ss in reqs:
            if g != grp:
                continue
            o = product.opt(**kw)
            keep.append((o, ns, a, ss))
            ids.append((L.ns_ext_batch_add(b, ns, len(ns), a, len(a), nsffi.C.byref(o), ss), ns, a, kw, ss))
        assert L.ns_ext_batch_flush(b) == len(ids)
        assert L.ns_ext_gpu_request_count() - before == len(ids)
        for rid, ns, a, kw, ss in ids:
            r = L.ns_ext_batch_get(b, rid).contents
            assert (r.nt_len, r.aa_len, r.score) == _oracle_ext(oracle, product, ns, a, kw, ss), (grp, len(ns), len(a))
    # several flushes in flight; ready/wait semantics (nasw.h:196-209)
    flights = []
    for grp in range(3):
        ids = []
        for g, ns, a, kw, ss in reqs:
            if g == grp:
                o = product.opt(**kw)
                keep.append(o)
                ids.append((L.ns_ext_batch_add(b, ns, len(ns), a, len(a), nsffi.C.byref(o), ss), ns, a, kw, ss))
        assert L.ns_ext_batch_flush(b) == len(ids)
        flights.append(ids)
    assert L.ns_ext_batch_flush(b) == 0
    for ids in flights:
        while not L.ns_ext_batch_ready(b, ids[0][0]):
            L.ns_ext_batch_wait(b)
        for rid, ns, a, kw, ss in ids:
            assert L.ns_ext_batch_ready(b, rid)
            r = L.ns_ext_batch_get(b, rid).contents
            assert (r.nt_len, r.aa_len, r.score) == _oracle_ext(oracle, product, ns, a, kw, ss)
    L.ns_ext_batch_wait(b)
    L.ns_ext_batch_destroy(b)


def test_ext_batch_over_limit_request_fails_alone(oracle, product):
    # one request past the device's 8192 padded columns, in the same option group as in-limit
    # ones: it alone gets INT32_MIN; the others are still computed, bit-exact
    L = product.lib
    reqs = [r for r in ref_style_requests() if r[0] == 0][:10]
    kw0 = reqs[0][3]
    rng = np.random.default_rng(5)
    big_a = bytes(rng.choice(list(b"ACDEFGHIKLMNPQRSTVWY"), 8200).astype(np.uint8))
    big_ns = bytes(rng.choice(list(b"ACGT"), 600).astype(np.uint8))
    b = L.ns_ext_batch_init(None)
    assert b, product.error()
    keep, ids = [], []
    for k, (g, ns, a, kw, ss) in enumerate(reqs):
        if k == 4:
            o = product.opt(**kw0)
            keep.append(o)
            ids.append((L.ns_ext_batch_add(b, big_ns, len(big_ns), big_a, len(big_a), nsffi.C.byref(o), None), None))
        o = product.opt(**kw)
        keep.append(o)
        ids.append((L.ns_ext_batch_add(b, ns, len(ns), a, len(a), nsffi.C.byref(o), ss), (ns, a, kw, ss)))
    assert L.ns_ext_batch_flush(b) == len(ids)
    for rid, x in ids:
        r = L.ns_ext_batch_get(b, rid).contents
        if x is None:
            assert (r.nt_len, r.aa_len, r.score) == (len(big_ns), len(big_a), -2 ** 31)
        else:
            assert (r.nt_len, r.aa_len, r.score) == _oracle_ext(oracle, product, *x)
    L.ns_ext_batch_destroy(b)


@pytest.mark.parametrize("seed,n,lo,hi", [(1, 384, 50, 1000), (2, 256, 20, 140), (3, 64, 900, 2000)])
def test_ext_plan_config1_batch(oracle, product, seed, n, lo, hi):
    from protlocus_amd import synth
    pairs = synth.extension_batch(n, seed=seed, min_len=lo, max_len=hi)
    rng = np.random.default_rng(seed)
    dirs = rng.integers(0, 2, size=n).astype(np.uint8)
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, product.opt(flag=NS_F_EXT_RIGHT), dirs=dirs)
    try:
        plan.run()
        got = plan.results()
        got2 = (plan.run(), plan.results())[1]   # a second run of the same plan is identical
        waves = plan.autotune(reps=1)            # the autotuner changes only the persistent grid
        got3 = (plan.run(), plan.results())[1]
    finally:
        plan.close()
    assert got == got2 == got3
    assert len(waves) == 3 and all(1 <= w <= 64 for w in waves), waves

    def one(k):
        d, p = pairs[k]
        kw = dict(flag=NS_F_EXT_LEFT if dirs[k] else NS_F_EXT_RIGHT)
        return _oracle_ext(oracle, product, d, p, kw, None)
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(one, range(n)))
    bad = [k for k in range(n) if got[k] != exp[k]]
    assert not bad, [(k, len(pairs[k][0]), len(pairs[k][1]), got[k], exp[k]) for k in bad[:5]]


@pytest.mark.parametrize("global_mode", [True, False])
def test_ext_plan_merged_queue_with_a_long_problem(oracle, product, global_mode):
    # a plan holding a long problem (rows x 128-column chunks >= 2^18: a 60-kb target, 600 aa)
    # and problems of 1-4 and 5-8 columns per lane runs its C <= 8 problems as one merged queue,
    # longest first, in one kernel; pairs, single problems of every column count, the C 9-12 group
    # beside it and a subset run must all equal the oracle
    from protlocus_amd import synth
    rng = np.random.default_rng(11 + global_mode)
    pairs = []
    p = synth.rand_protein(rng, 600)
    if global_mode:
        pairs.append((synth.rand_dna(rng, 60000), p))
    else:   # the protein's gene with a 60-kb intron: the extension crosses it
        pairs.append((synth.gene_of(rng, p, 0.1, 1, lambda: 60000, 0, 200), p))
    for lo, hi, m in [(1, 250, 40), (257, 512, 20), (520, 1000, 20), (1030, 1500, 4)]:
        for _ in range(m):
            al = int(rng.integers(lo, hi + 1))
            q = synth.rand_protein(rng, al)
            if rng.integers(0, 3) == 0:
                d = synth.rand_dna(rng, int(rng.integers(3, 3 * al + 400)))
            else:
                d = synth.gene_of(rng, q, 0.2, int(rng.integers(0, 3)), lambda: int(rng.integers(40, 400)), 1, 200)
            pairs.append((d, q))
    assert len(pairs[0][0]) * ((600 + 127) // 128) >= 1 << 18
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    flag = 0 if global_mode else NS_F_EXT_RIGHT
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, product.opt(flag=flag))
    try:
        plan.run()
        got = plan.results()
        keep = np.zeros(len(pairs), np.uint8)
        keep[::2] = 1
        assert product.lib.nsh_ext_plan_run_subset(plan.p, bytes(keep)), product.error()
        assert product.lib.nsh_ext_plan_sync(plan.p)
        sub = plan.results()
    finally:
        plan.close()

    def one(k):
        d, q = pairs[k]
        if global_mode:
            return (len(d), len(q), oracle.global_dp(d, q, oracle.opt(flag=0), mode=16)[2])
        return _oracle_ext(oracle, product, d, q, dict(flag=NS_F_EXT_RIGHT), None)
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(one, range(len(pairs))))
    bad = [k for k in range(len(pairs)) if got[k] != exp[k]]
    assert not bad, [(k, len(pairs[k][0]), len(pairs[k][1]), got[k], exp[k]) for k in bad[:5]]
    bad = [k for k in range(0, len(pairs), 2) if sub[k] != exp[k]]
    assert not bad, [(k, sub[k], exp[k]) for k in bad[:5]]


@pytest.mark.parametrize("xdrop,global_mode", [(20, False), (60, False), (100, True)])
def test_ext_plan_two_problems_per_wave(oracle, product, xdrop, global_mode):
    # problems of at most 256 padded columns share a wave two at a time (one per 16-bit half):
    # lengths 1-250 aa, pairs whose partners stop at very different rows (small x-drop, targets of
    # 3 to 3,000 nt), both directions in one plan, or the score-only global mode; every result must
    # equal the oracle's, like the one-problem-per-wave path
    from protlocus_amd import synth
    rng = np.random.default_rng(xdrop)
    pairs = []
    for k in range(300):
        al = int(rng.integers(1, 251))
        p = synth.rand_protein(rng, al)
        if k % 3 == 0:   # random DNA of 3 to 3,000 nt: an extension stops early or runs to the end
            d = synth.rand_dna(rng, int(rng.integers(3, 3000)))
        else:            # the protein's own gene (introns, frameshifts, 200-nt flanks)
            d = synth.gene_of(rng, p, 0.2, int(rng.integers(0, 3)), lambda: int(rng.integers(40, 400)), 1, 200)
        pairs.append((d, p))
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    if global_mode:
        flag, dirs = 0, None
    else:
        flag, dirs = NS_F_EXT_RIGHT, rng.integers(0, 2, size=len(pairs)).astype(np.uint8)
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, product.opt(flag=flag, xdrop=xdrop), dirs=dirs)
    try:
        plan.run()
        got = plan.results()
        keep = np.zeros(len(pairs), np.uint8)
        keep[::2] = 1           # a subset run unpacks the pairs into single problems
        assert product.lib.nsh_ext_plan_run_subset(plan.p, bytes(keep)), product.error()
        assert product.lib.nsh_ext_plan_sync(plan.p)
        sub = plan.results()
    finally:
        plan.close()

    def one(k):
        d, p = pairs[k]
        if global_mode:   # ns_global_gs16b without CIGAR or EXT flags: (nl, al, H(last row, al-1))
            return (len(d), len(p), oracle.global_dp(d, p, oracle.opt(flag=0, xdrop=xdrop), mode=16)[2])
        kw = dict(flag=NS_F_EXT_LEFT if dirs[k] else NS_F_EXT_RIGHT, xdrop=xdrop)
        return _oracle_ext(oracle, product, d, p, kw, None)
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(one, range(len(pairs))))
    bad = [k for k in range(len(pairs)) if got[k] != exp[k]]
    assert not bad, [(k, len(pairs[k][0]), len(pairs[k][1]), got[k], exp[k]) for k in bad[:5]]
    bad = [k for k in range(0, len(pairs), 2) if sub[k] != exp[k]]
    assert not bad, [(k, sub[k], exp[k]) for k in bad[:5]]


def test_ext_plan_graph_replays_match_eager_runs(oracle, product):
    # the plan's run captured as one hipGraph and replayed: every replay equals the eager run's
    # results, with other eager work and synchronisations between the replays; a second plan's
    # graph is independent of the first; after the autotuner changes the grids the graph is
    # captured again
    from protlocus_amd import synth
    pairs = synth.extension_batch(192, seed=8, min_len=20, max_len=1200)
    rng = np.random.default_rng(8)
    dirs = rng.integers(0, 2, size=len(pairs)).astype(np.uint8)
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    other = synth.extension_batch(64, seed=9, min_len=50, max_len=400)
    ons, ons_off, onl, oas, oas_off, oal = synth.pack_batch(other)
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, product.opt(flag=NS_F_EXT_RIGHT), dirs=dirs)
    plan2 = product.ext_plan(ons, ons_off, onl, oas, oas_off, oal, product.opt(flag=NS_F_EXT_LEFT))
    try:
        plan.run()
        eager = plan.results()
        plan2.run()
        eager2 = plan2.results()
        for rep in range(4):
            plan.run(graph=True)
            assert plan.results() == eager, rep
            assert plan.ms(1) > 0
            plan2.run()                     # eager work of another plan between the replays
            assert plan2.results() == eager2, rep
            plan2.run(graph=True)
            assert plan2.results() == eager2, rep
        plan.autotune(reps=1)
        for rep in range(2):
            plan.run(graph=True)
            assert plan.results() == eager, ("after autotune", rep)
    finally:
        plan.close()
        plan2.close()

    def one(k):
        d, p = pairs[k]
        return _oracle_ext(oracle, product, d, p, dict(flag=NS_F_EXT_LEFT if dirs[k] else NS_F_EXT_RIGHT), None)
    with cf.ThreadPoolExecutor(8) as ex:
        exp = list(ex.map(one, range(len(pairs))))
    assert eager == exp
