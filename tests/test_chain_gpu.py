"""GPU parity of the batched anchor chaining nsh_chain_batch against the reference's mp_chain
(chain.c:7-219, as map.c:611-633 calls it): every query's chains (score << 32 | count) and their
anchors must be identical, in the same order. The checks use the golden records of the compiled
reference, grouped into batches by parameter set, and random batches (empty queries among them)
against the pinned restatement tests/chain_ref.py."""
import gzip
import json
import os
from collections import defaultdict

import numpy as np
import pytest

import chain_ref
from cases import anchor_chain_case, unpack_keys

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def test_chain_golden(product):
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "anchor_chain_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    groups = defaultdict(list)
    for k, c in enumerate(cases):
        groups[tuple(c["params"])].append(k)
    for params, ks in groups.items():
        print("params", params, "queries", len(ks), flush=True)
        got = capi.chain_batch([unpack_keys(cases[k]["anchors"]) for k in ks], *params)
        for k, (u, a) in zip(ks, got):
            assert u.tolist() == unpack_keys(cases[k]["u"]).tolist(), k
            assert a.tolist() == unpack_keys(cases[k]["a"]).tolist(), k


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_chain_random(product, seed):
    from protlocus_amd import capi
    rng = np.random.default_rng(seed)
    k0 = int(rng.integers(0, 1000))
    base = anchor_chain_case(rng, k0, n_max=10)
    params = base["params"]
    qs = []
    for j in range(int(rng.integers(20, 70))):
        c = anchor_chain_case(rng, k0 + 60 * j, n_max=int(rng.choice([5, 100, 900])))   # same kmer, bbit, params
        qs.append(c["anchors"] if rng.random() > 0.1 else np.zeros(0, np.uint64))
    print("seed", seed, "params", params, "queries", len(qs), "anchors", sum(len(q) for q in qs), flush=True)
    got = capi.chain_batch(qs, *params)
    for j, (q, (u, a)) in enumerate(zip(qs, got)):
        want = chain_ref.chain([int(x) for x in q], *params)
        assert (u.tolist(), a.tolist()) == want, (seed, j)


def test_chain_long_diagonals(product):
    """one query of 30,000 anchors in long co-linear runs and a perfect diagonal"""
    from protlocus_amd import capi
    n = 20000
    diag = np.array([(3 * i) << 32 | i for i in range(n)], np.uint64)
    (u, a), = capi.chain_batch([diag], 5000, 5000, 500, 25, 5000, 2, 0, 0.75, 1, 1, 0)
    assert u.tolist() == [n << 32 | n] and a.tolist() == diag.tolist()
    rng = np.random.default_rng(11)
    c = anchor_chain_case(rng, 1, n_max=1)
    xs = []
    for r in range(300):
        t, q = int(rng.integers(0, 5000)), int(rng.integers(0, 3000))
        for i in range(100):
            q += int(rng.integers(1, 4))
            t += int(rng.integers(0, 3))
            xs.append(t << 32 | q)
    big = np.sort(np.array(xs, np.uint64))
    for params in ((256, 256, 256, 25, 1000000, 2, 0, 0.75, 1, 4, 6), c["params"]):
        print("params", params, flush=True)
        (u, a), = capi.chain_batch([big], *params)
        assert (u.tolist(), a.tolist()) == chain_ref.chain([int(x) for x in big], *params)


def test_chain_rejects_bad_arguments(product):
    from protlocus_amd.capi import lib
    off = np.array([0, 2, 1], np.int64)
    a = np.zeros(2, np.uint64)
    n_u = np.zeros(2, np.int32)
    out = np.zeros(2, np.uint64)
    assert lib().nsh_chain_batch(2, off.ctypes.data, a.ctypes.data, 1, 1, 1, 1, 1, 1, 0, 0.75, 1, 3, 8,
                                 n_u.ctypes.data, out.ctypes.data, out.ctypes.data) < 0
    off = np.array([0, 1, 2], np.int64)
    assert lib().nsh_chain_batch(2, off.ctypes.data, a.ctypes.data, 1, 1, 1, 1, 1, 1, 0, 0.75, 1, 0, 8,
                                 n_u.ctypes.data, out.ctypes.data, out.ctypes.data) < 0
    assert lib().nsh_chain_batch(0, None, None, 1, 1, 1, 1, 1, 1, 0, 0.75, 1, 3, 8, None, None, None) == 0


def _coarse_ref(anchors, opt, kmer, bbit):
    """map.c:611-633 on one query's sorted anchors, through the restatement"""
    a = [int(x) for x in anchors]
    if opt.pre_chain and opt.is_spliced:
        w = 1 << bbit
        _, a = chain_ref.chain(a, w, w, w, opt.max_skip, opt.max_iter, 2, 0, opt.chn_coef_log, opt.is_spliced, kmer,
                               bbit)
        a = sorted(a)
    return chain_ref.chain(a, opt.max_dist_x, opt.max_dist_y, opt.bw, opt.max_skip, opt.max_iter, opt.min_cnt,
                           opt.min_sc, opt.chn_coef_log, opt.is_spliced, kmer, bbit)


def test_index_chain_golden(product):
    """nsh_index_chain (anchors, pre-chain, sort, chain on the device) against the reference's
    anchors of the golden records chained by the restatement, for three option sets"""
    import base64
    import zlib
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "anchor_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    opts = [capi.ChainOpt(), capi.ChainOpt(pre_chain=0, min_cnt=2, max_dist_x=5000, bw=500),
            capi.ChainOpt(is_spliced=0, max_dist_x=20000, max_dist_y=2000, bw=1000, min_sc=5)]
    for k, c in enumerate(cases):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        ix = capi.Index(g, c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"])
        try:
            qs = [q.encode() for q in c["queries"]]
            for oi, opt in enumerate(opts):
                if k == 3 and oi:   # the repeat-rich genome: its restatement side takes seconds per set
                    continue
                print("case", k, "options", oi, flush=True)
                got = ix.chain(qs, c["max_occ"], opt)
                for j, (q, (u, a)) in enumerate(zip(c["anchors"], got)):
                    want = _coarse_ref(unpack_keys(q), opt, c["kmer"], c["bbit"])
                    assert (u.tolist(), a.tolist()) == want, (k, oi, j)
        finally:
            ix.close()


def test_chain_large_staged(product):
    """1,000 queries of 9,001 anchors, each on a perfect diagonal: 72 MB each way, more than one
    64-MB staging chunk of the host copies; every query gives one chain of all its anchors"""
    from protlocus_amd import capi
    m, nq = 9001, 1000
    i = np.arange(m, dtype=np.uint64)
    qs = [((np.uint64(3) * i + np.uint64(q * 50000)) << np.uint64(32)) | (i + np.uint64(q)) for q in range(nq)]
    print("anchors", m * nq, flush=True)
    got = capi.chain_batch(qs, 100, 100, 100, 25, 5000, 2, 0, 0.75, 1, 1, 0)
    for q, (u, a) in enumerate(got):
        assert u.tolist() == [m << 32 | m], q
        assert np.array_equal(a, qs[q]), q


def test_copies_without_pinned_staging(product, monkeypatch):
    """NSH_NO_PINNED_STAGING=1 makes the large host copies plain copies on the entry point's stream
    followed by a wait on it: the chain batch of 72 MB each way and the index of a 40-Mb genome
    must come out as they do through the pinned staging buffers"""
    from protlocus_amd import capi
    m, nq = 9001, 1000
    i = np.arange(m, dtype=np.uint64)
    qs = [((np.uint64(3) * i + np.uint64(q * 50000)) << np.uint64(32)) | (i + np.uint64(q)) for q in range(nq)]
    rng = np.random.default_rng(8)
    g = rng.integers(0, 4, size=40_000_000, dtype=np.uint8)
    staged = capi.index_build(g, [0, 25_000_000], [25_000_000, 15_000_000], 30, 6, 1, 8)
    monkeypatch.setenv("NSH_NO_PINNED_STAGING", "1")
    print("chain batch, unstaged", flush=True)
    got = capi.chain_batch(qs, 100, 100, 100, 25, 5000, 2, 0, 0.75, 1, 1, 0)
    for q, (u, a) in enumerate(got):
        assert u.tolist() == [m << 32 | m], q
        assert np.array_equal(a, qs[q]), q
    print("index build, unstaged", flush=True)
    plain = capi.index_build(g, [0, 25_000_000], [25_000_000, 15_000_000], 30, 6, 1, 8)
    for x, y in zip(staged, plain):
        assert np.array_equal(x, y)


def test_index_chain_edges(product):
    """nsh_index_chain: a null index or options are refused; no queries give no chains; empty and
    3-residue queries give none either"""
    import ctypes as C
    from protlocus_amd import capi
    L = capi.lib()
    uoff, coff = np.zeros(3, np.int64), np.zeros(3, np.int64)
    opt = capi.ChainOpt()
    assert L.nsh_index_chain(None, 0, b"", None, None, 20000, C.byref(opt), uoff.ctypes.data, None, 0,
                             coff.ctypes.data, None, 0) < 0
    rng = np.random.default_rng(8)
    g = rng.integers(0, 4, 30000).astype(np.uint8)
    ix = capi.Index(g, [0], [len(g)])
    try:
        assert L.nsh_index_chain(ix.p, 0, b"", None, None, 20000, None, uoff.ctypes.data, None, 0,
                                 coff.ctypes.data, None, 0) < 0
        assert L.nsh_index_chain(ix.p, 0, b"", None, None, 20000, C.byref(opt), uoff.ctypes.data, None, 0,
                                 coff.ctypes.data, None, 0) == 0
        assert uoff[0] == 0 and coff[0] == 0
        got = ix.chain([b"", b"MKV"])
        assert [(len(u), len(a)) for u, a in got] == [(0, 0), (0, 0)]
    finally:
        ix.close()
