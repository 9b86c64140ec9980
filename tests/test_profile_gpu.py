"""GPU parity of the profile and gapless paths against the oracle (bit-exact):
  * profile-HMM Viterbi (nsh_hmm_*, behind mp_hmm_dp_score_gpu*): the golden profiles and
    regions written from the compiled reference, and random profiles with M at every lane
    boundary up to 1024 against regions with N bases, stop codons and lengths 0..3000
    (test/test_gpu.c:221-290 compares the GPU with mp_hmm_dp_score the same way);
  * gapless X-drop candidate scores (ns_gpu_gapless_score_batch) against the oracle's
    restatement of nasw-cuda.cu:1151-1295, forward and reverse strands, k-mer 1..8, block
    widths 1..300, x-drop 0..60, anchors out of range and more than 32 anchors;
  * the ungapped PSSM diagonal (nsh_pssm_ungapped_batch) against prescreen.c:334-344.
"""
import ctypes as C
import gzip
import json
import os

import numpy as np
import pytest

import nsffi

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _parse(oracle, text):
    h = nsffi.OrHmm()
    assert oracle.lib.or_hmm_parse(text.encode(), C.byref(h), None) == 1
    M = h.len
    emit = np.ctypeslib.as_array(h.emit, shape=(M * 20,)).copy()
    trans = np.ctypeslib.as_array(h.trans, shape=(M * 7,)).copy()
    return h, M, emit, trans


def _gpu_scores(product, M, emit, trans, regions):
    L = product.lib
    g = L.nsh_hmm_create(M, emit.ctypes.data, trans.ctypes.data)
    assert g, product.error()
    try:
        nt = b"".join(regions)
        off = np.cumsum([0] + [len(r) for r in regions[:-1]]).astype(np.int32)
        ln = np.array([len(r) for r in regions], dtype=np.int32)
        out = np.zeros(len(regions), dtype=np.int32)
        assert L.nsh_hmm_score_batch(g, len(regions), nt, len(nt), off.ctypes.data, ln.ctypes.data, out.ctypes.data), \
            product.error()
        return out.tolist()
    finally:
        L.nsh_hmm_destroy(g)


def test_hmm_golden(oracle, product):
    with gzip.open(os.path.join(GOLD, "hmm_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for case in cases:
        h, M, emit, trans = _parse(oracle, case["hmm"])
        oracle.lib.or_hmm_free(C.byref(h))
        regions = [bytes.fromhex(r) for r in case["regions"]]
        assert _gpu_scores(product, M, emit, trans, regions) == case["scores"], M


@pytest.mark.parametrize("M", [1, 2, 63, 64, 65, 127, 128, 129, 200, 333, 511, 512, 513, 800, 1000, 1024])
def test_hmm_random(oracle, product, M):
    from protlocus_amd import synth
    rng = np.random.default_rng(M)
    em, tr = synth.hmm_params(rng, M)
    h, M2, emit, trans = _parse(oracle, synth.hmm_text(em, tr))
    assert M2 == M
    regions = []
    for k in range(24):
        if k % 3 == 0:   # a profile-sampled protein, reverse-translated, in frame
            prot = synth.hmm_sample_protein(rng, em, tr)
            cod = np.frombuffer(synth.reverse_translate(rng, prot.encode()), dtype=np.uint8)
            lut = np.full(256, 4, dtype=np.uint8)
            lut[list(b"ACGT")] = [0, 1, 2, 3]
            nt = np.concatenate([rng.integers(0, 4, size=3 * int(rng.integers(0, 40))), lut[cod],
                                 rng.integers(0, 4, size=int(rng.integers(0, 100)))]).astype(np.uint8)
        else:
            nt = rng.integers(0, 5 if k % 3 == 1 else 4, size=int(rng.integers(0, 3000))).astype(np.uint8)
        regions.append(nt.tobytes())
    exp = [oracle.lib.or_hmm_viterbi(C.byref(h), r, len(r)) for r in regions]
    oracle.lib.or_hmm_free(C.byref(h))
    assert _gpu_scores(product, M, emit, trans, regions) == exp


def test_hmm_rejects_long_profiles(product):
    emit = np.zeros(1025 * 20, dtype=np.int16)
    trans = np.zeros(1025 * 7, dtype=np.int16)
    assert not product.lib.nsh_hmm_create(1025, emit.ctypes.data, trans.ctypes.data)


@pytest.mark.parametrize("seed,kmer,bw,xdrop", [(1, 5, 256, 30), (2, 3, 64, 0), (3, 7, 300, 60), (4, 1, 1, 10),
                                                 (5, 8, 100, 25), (6, 4, 17, 5), (7, 2, 32, 1000)])
def test_gapless_batch(oracle, product, seed, kmer, bw, xdrop):
    from cases import gapless_batch, pack4
    rng = np.random.default_rng(seed)
    codon = oracle.codon
    g, queries, cands, anchors = gapless_batch(rng, 300, codon, kmer, bw)
    pk = pack4(g)
    mat = np.array(oracle.blosum62, dtype=np.int8)
    CA = (nsffi.GaplessCand * len(cands))(*cands)
    AN = (nsffi.GaplessAnchor * max(1, len(anchors)))(*anchors)
    out = np.zeros(len(cands), dtype=np.int32)
    L = product.lib
    # the genome is cached by (pointer, length) like the reference (nasw-cuda.cu:1349-1356); a new
    # genome object may reuse a freed one's address, so the caller resets the cache (nasw.h:178)
    L.ns_gpu_gapless_reset_target()
    ok = L.ns_gpu_gapless_score_batch(pk, len(g), queries, len(queries), CA, len(cands), AN, max(1, len(anchors)),
                                      bw, kmer, xdrop, mat.ctypes.data, 22, codon, out.ctypes.data)
    assert ok, product.error()
    exp = [oracle.lib.or_gapless_score(pk, queries, C.byref(c), AN, bw, kmer, xdrop, mat.ctypes.data, 22, codon)
           for c in cands]
    bad = [(i, int(out[i]), exp[i]) for i in range(len(cands)) if out[i] != exp[i]]
    assert not bad, bad[:5]
    assert any(e > 100 for e in exp) and any(e == -2 ** 31 for e in exp)   # planted hits and empty candidates
    prof = nsffi.GaplessProfile()
    L.ns_gpu_gapless_profile(C.byref(prof))
    assert prof.batches >= 1 and prof.candidates >= len(cands)


def test_gapless_closed_form_rules(oracle, product):
    """the constructed candidates whose scores are fixed by construction (cases.gapless_closed_forms:
    proration of an overlapped segment, the x-drop rule for later segments, a covered segment
    dropped, equal scores in anchor order, the strict x-drop stop of the walks), on the GPU"""
    from cases import pack4
    from test_oracle_cpu import _closed_form_scores
    L = product.lib

    def gpu(g, q, cs, ans, bw, kmer, xdrop, mat):
        CA = (nsffi.GaplessCand * len(cs))(*cs)
        AN = (nsffi.GaplessAnchor * len(ans))(*ans)
        out = np.zeros(len(cs), dtype=np.int32)
        L.ns_gpu_gapless_reset_target()
        ok = L.ns_gpu_gapless_score_batch(pack4(g), len(g), bytes(q), len(q), CA, len(cs), AN, len(ans), bw, kmer,
                                          xdrop, mat.ctypes.data, 22, oracle.codon, out.ctypes.data)
        assert ok, product.error()
        return out.tolist()
    for name, got, exp in _closed_form_scores(gpu, oracle.codon):
        assert got == exp, name


def test_gapless_rejects_out_of_range(oracle, product):
    from cases import pack4
    g = np.zeros(100, dtype=np.uint8)
    q = bytes(10)
    c = nsffi.GaplessCand(target_offset=50, target_length=60, query_offset=0, query_length=10, anchor_offset=0,
                          anchor_count=1, reverse=0)
    a = nsffi.GaplessAnchor(target_block=0, query_end=5, reserved=0)
    out = np.zeros(1, dtype=np.int32)
    mat = np.array(oracle.blosum62, dtype=np.int8)
    assert not product.lib.ns_gpu_gapless_score_batch(pack4(g), 100, q, 10, C.byref(c), 1, C.byref(a), 1, 16, 3, 10,
                                                      mat.ctypes.data, 22, oracle.codon, out.ctypes.data)


def test_pssm_ungapped_batch(oracle, product):
    rng = np.random.default_rng(17)
    for L in (1, 7, 64, 300, 1500):
        pssm = rng.integers(-8, 9, size=(L, 22)).astype(np.int8)
        orfs = [rng.integers(0, 22, size=int(rng.integers(1, 700))).astype(np.uint8) for _ in range(50)]
        aa = b"".join(o.tobytes() for o in orfs)
        offs = np.cumsum([0] + [len(o) for o in orfs[:-1]]).astype(np.int64)
        items = [(k, int(rng.integers(-L + 1, len(orfs[k])))) for k in range(50) for _ in range(8)]
        io = np.array([offs[k] for k, _ in items], dtype=np.int64)
        il = np.array([len(orfs[k]) for k, _ in items], dtype=np.int32)
        idg = np.array([d for _, d in items], dtype=np.int32)
        n = len(items)
        sc, qs, qe = (np.zeros(n, dtype=np.int32) for _ in range(3))
        assert product.lib.nsh_pssm_ungapped_batch(pssm.ctypes.data, L, aa, len(aa), n, io.ctypes.data,
                                                   il.ctypes.data, idg.ctypes.data, sc.ctypes.data, qs.ctypes.data,
                                                   qe.ctypes.data), product.error()
        for i, (k, d) in enumerate(items):
            a, b = C.c_int32(), C.c_int32()
            e = oracle.lib.or_pssm_ungapped(pssm.ctypes.data, L, orfs[k].tobytes(), len(orfs[k]), d, C.byref(a),
                                            C.byref(b))
            assert (sc[i], qs[i], qe[i]) == (e, a.value, b.value), (L, i, k, d)


def _tabs(oracle):
    L = oracle.lib
    return list((C.c_uint8 * 64).in_dll(L, "or_codon")), list((C.c_uint8 * 256).in_dll(L, "or_aa20"))[ord("X")]


def test_chain_pssm_evidence_golden(oracle, product):
    """nsh_chain_pssm_ev_batch against mp_chain_pssm_ev / mp_diag_seg of the compiled reference
    (golden records), all chains of a parameter set in one batch"""
    import base64
    import pssm_ref
    from protlocus_amd import capi
    codon, ax = _tabs(oracle)
    with gzip.open(os.path.join(HERE, "golden", "pssm_chain_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    groups = {}
    for c in cases:
        groups.setdefault((c["kmer"], c["min_seg"], c["xdrop"]), []).append(c)
    for (kmer, min_seg, xdrop), cs in groups.items():
        chains = [(np.frombuffer(base64.b64decode(c["nt"]), dtype=np.uint8),
                   np.frombuffer(base64.b64decode(c["pssm"]), dtype=np.int8).reshape(-1, 22), c["anchors"]) for c in cs]
        ev, per = capi.chain_pssm_ev(chains, kmer, min_seg, xdrop)
        assert ev == [c["ev"] for c in cs]
        for c, ch, pa in zip(cs, chains, per):
            _, exp_per = pssm_ref.chain_pssm_ev(ch[0].tolist(), ch[1].ravel().tolist(), c["qlen"], kmer, c["anchors"],
                                                min_seg, xdrop, codon, ax)
            assert [p is None for p in pa] == [p is None for p in exp_per]
            for p, seg in zip(pa, c["diag_seg"]):
                if p is not None:
                    assert list(p) == seg


@pytest.mark.parametrize("seed", [1, 2])
def test_chain_pssm_evidence_random(oracle, product, seed):
    import pssm_ref
    from cases import chain_case
    from protlocus_amd import capi
    codon, ax = _tabs(oracle)
    rng = np.random.default_rng(seed)
    for kmer, min_seg, xdrop in ((3, 15, 20), (5, 0, 3), (4, -1000, 40)):
        cs = [chain_case(rng, 3 * i + kmer - 3) for i in range(200)]    # chain_case(k) has kmer 3 + k % 3
        assert all(c["kmer"] == kmer for c in cs)
        chains = [(c["nt"], c["pssm"], c["anchors"]) for c in cs]
        chains.append((np.zeros(0, np.uint8), np.zeros((8, 22), np.int8), [2 << 32 | 4]))   # empty region
        chains.append((cs[0]["nt"], cs[0]["pssm"], []))                                      # no anchors
        ev, per = capi.chain_pssm_ev(chains, kmer, min_seg, xdrop)
        for k, (nt, pssm, an) in enumerate(chains):
            exp, exp_per = pssm_ref.chain_pssm_ev(nt.tolist(), pssm.ravel().tolist(), len(pssm), kmer, an, min_seg,
                                                  xdrop, codon, ax)
            assert (ev[k], per[k]) == (exp, exp_per), k


def test_orf_scan_golden(oracle, product):
    import base64
    from cases import revcomp_codes
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "orf_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for c in cases:
        seq = np.frombuffer(base64.b64decode(c["seq"]), dtype=np.uint8)
        for rev in (0, 1):
            orfs, aa = capi.orf_scan(seq if rev == 0 else revcomp_codes(seq), c["min_orf"])
            assert [list(x) for x in orfs] == c["orfs"][rev]
            assert bytes(aa) == base64.b64decode(c["aa"][rev])


def test_orf_scan_random_and_edges(oracle, product):
    """long sequences (several chunks per frame), N runs, stop-free stretches, tiny inputs"""
    import pssm_ref
    from cases import nt4_codes, orf_case
    from protlocus_amd import capi, synth
    codon, _ = _tabs(oracle)
    rng = np.random.default_rng(17)
    seqs = [np.zeros(0, np.uint8), np.array([0, 1], np.uint8), np.array([0, 3, 2], np.uint8),
            np.full(50, 4, np.uint8)]
    prot = synth.rand_protein(rng, 9000)                        # one stop-free frame across chunks
    seqs.append(nt4_codes(synth.reverse_translate(rng, prot)))
    for _ in range(3):
        seqs.append(np.concatenate([orf_case(rng, int(rng.integers(0, 8)))["seq"] for _ in range(120)]))
    for s in seqs:
        for m in (1, 30):
            exp = pssm_ref.orf_walk(s.tolist(), m, codon)
            assert capi.orf_scan(s, m) == exp, (len(s), m)


def test_ps_gapped_golden(product):
    """nsh_ps_gapped_batch against ps_gapped of the compiled reference (golden records)"""
    import base64
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "ps_gapped_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for c in cases:
        pssm = np.frombuffer(base64.b64decode(c["pssm"]), dtype=np.int8).reshape(-1, 22)
        orfs = [np.frombuffer(base64.b64decode(o), dtype=np.uint8) for o in c["orfs"]]
        assert capi.ps_gapped_batch(pssm, orfs, c["diags"], c["band"], c["go"], c["ge"]) == c["sw"]


def test_ps_gapped_random(product):
    """larger batches: every band width up to the 1024-column limit, full matrices up to 1024 aa"""
    import pssm_ref
    from cases import gapped_case
    from protlocus_amd import capi
    rng = np.random.default_rng(23)
    for k in range(12):
        c = gapped_case(rng, k)
        orfs, diags = list(c["orfs"]), list(c["diags"])
        for _ in range(40):
            n = int(rng.integers(0, 1025 if c["band"] <= 0 else 700))
            orfs.append(rng.integers(0, 20, size=n).astype(np.uint8))
            diags.append(int(rng.integers(-len(c["pssm"]) - 40, n + 40)))
        band = [32, 100, 511, 0, 7, 64][k % 6]
        got = capi.ps_gapped_batch(c["pssm"], orfs, diags, band, c["go"], c["ge"])
        flat = c["pssm"].ravel().tolist()
        for o, d, g in zip(orfs, diags, got):
            assert g == pssm_ref.ps_gapped(flat, len(c["pssm"]), o.tolist(), d, band, c["go"], c["ge"]), (k, len(o), d)


def test_prescreen_search_golden(product):
    """nsh_prescreen_build + nsh_prescreen_search against the reference's mp_prescreen_build +
    ps_query_worker (golden records): the ORF database size and the candidate list, field by
    field (E-values included) and in order; spaced 5-, 6- and 7-mers, banded and full gapped
    stages, E-value cuts, k-mer caps and max_seqs truncation"""
    import base64
    import zlib
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "prescreen_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for k, c in enumerate(cases):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        p = c["params"]
        print("case", k, "genome", len(g), "kmer", p["kmer"], flush=True)
        ps = capi.ProfileSearch(g, c["ctg_off"], c["ctg_len"], kmer=p["kmer"], min_orf=p["min_orf"])
        try:
            assert [ps.size(0), ps.size(1)] == c["size"], k
            pssm = np.frombuffer(base64.b64decode(c["pssm"]), dtype=np.int8).reshape(-1, 22)
            got = ps.search(pssm, c["thr"], kmer_max=p["kmer_max"], min_ungapped=p["min_ungapped"],
                            max_seqs=p["max_seqs"], band=p["band"], max_evalue=p["max_evalue"])
            assert [list(x) for x in got] == c["cands"], (k, got, c["cands"])
        finally:
            ps.close()


def test_prescreen_search_batch_golden(product):
    """nsh_prescreen_search_batch against the reference's per-profile ps_query_worker lists over one
    index (golden records): every profile's list in order, profiles shorter than the pattern span,
    profiles without candidates, the k-mer cap, the search counters; each profile searched alone
    agrees too"""
    import base64
    import zlib
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "prescreen_batch_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for k, c in enumerate(cases):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        p = c["params"]
        print("case", k, "genome", len(g), "profiles", len(c["pssms"]), "kmer", p["kmer"], flush=True)
        pssms = [np.frombuffer(base64.b64decode(x), dtype=np.int8).reshape(-1, 22) for x in c["pssms"]]
        kw = dict(kmer_max=p["kmer_max"], min_ungapped=p["min_ungapped"], max_seqs=p["max_seqs"], band=p["band"],
                  max_evalue=p["max_evalue"])
        ps = capi.ProfileSearch(g, c["ctg_off"], c["ctg_len"], kmer=p["kmer"], min_orf=p["min_orf"])
        try:
            got = ps.search_batch(pssms, c["thr"], **kw)
            assert [[list(x) for x in r] for r in got] == c["cands"], k
            # the worker's counters, summed over the batch (prescreen.c:503-525)
            assert list(ps.stats().values()) == [sum(s[i] for s in c["stats"]) for i in range(7)], k
            for j, m in enumerate(pssms):
                assert [list(x) for x in ps.search(m, c["thr"], **kw)] == c["cands"][j], (k, j)
            assert ps.search_batch([], c["thr"], **kw) == []
        finally:
            ps.close()
