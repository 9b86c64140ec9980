"""The PSSM diagonal evidence and six-frame ORF restatements (tests/pssm_ref.py) against the
compiled reference: its golden records (tests/golden/pssm_chain_cases.json.gz, orf_cases.json.gz),
fresh random cases when oracle/_ref is built, and closed forms that do not depend on it."""
import base64
import ctypes as C
import gzip
import json
import os

import numpy as np
import pytest

import nsffi
import pssm_ref
from cases import chain_case, nt4_codes, orf_case, revcomp_codes

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    with gzip.open(os.path.join(HERE, "golden", name), "rt") as f:
        return json.load(f)


@pytest.fixture(scope="module")
def tabs(oracle):
    L = oracle.lib
    codon = list((C.c_uint8 * 64).in_dll(L, "or_codon"))
    aa20 = list((C.c_uint8 * 256).in_dll(L, "or_aa20"))
    return codon, aa20[ord("X")]


def test_chain_evidence_golden(tabs):
    codon, ax = tabs
    n_dup = 0
    for c in _load("pssm_chain_cases.json.gz"):
        nt = list(base64.b64decode(c["nt"]))
        pssm = np.frombuffer(base64.b64decode(c["pssm"]), dtype=np.int8).tolist()
        ev, per = pssm_ref.chain_pssm_ev(nt, pssm, c["qlen"], c["kmer"], c["anchors"], c["min_seg"], c["xdrop"], codon, ax)
        assert ev == c["ev"]
        for av, exp in zip(c["anchors"], c["diag_seg"]):
            assert list(pssm_ref.diag_seg(nt, pssm, c["qlen"], c["kmer"], av >> 32, av & 0xffffffff, c["xdrop"], codon,
                                          ax)) == exp
        n_dup += sum(p is None for p in per)
    assert n_dup > 10          # the duplicate-diagonal rule is exercised


def test_orf_walk_golden(tabs):
    codon, _ = tabs
    for c in _load("orf_cases.json.gz"):
        seq = np.frombuffer(base64.b64decode(c["seq"]), dtype=np.uint8)
        for rev in (0, 1):
            s = seq if rev == 0 else revcomp_codes(seq)
            orfs, aa = pssm_ref.orf_walk(s.tolist(), c["min_orf"], codon)
            assert [list(x) for x in orfs] == c["orfs"][rev]
            assert bytes(aa) == base64.b64decode(c["aa"][rev])


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_restatements_match_reference_live(tabs):
    codon, ax = tabs
    ref = nsffi.Reference()
    rng = np.random.default_rng(99)
    for k in range(120):
        c = chain_case(rng, k)
        flat = c["pssm"].ravel().tolist()
        ev, _ = pssm_ref.chain_pssm_ev(c["nt"].tolist(), flat, c["qlen"], c["kmer"], c["anchors"], c["min_seg"],
                                       c["xdrop"], codon, ax)
        assert ev == ref.chain_pssm_ev(c["nt"], c["pssm"], c["qlen"], c["kmer"], c["anchors"], c["min_seg"], c["xdrop"])
    for k in range(60):
        o = orf_case(rng, k)
        for rev in (0, 1):
            s = o["seq"] if rev == 0 else revcomp_codes(o["seq"])
            assert pssm_ref.orf_walk(s.tolist(), o["min_orf"], codon) == ref.orfs(o["seq"], rev, o["min_orf"])


def test_closed_forms(tabs):
    """without an x-drop stop, a diagonal's segment is the best forward prefix sum plus the best
    backward one; ORFs are maximal stop- and N-free codon runs of their frame"""
    from protlocus_amd import synth
    codon, ax = tabs
    rng = np.random.default_rng(7)
    for _ in range(40):
        qlen, kmer = int(rng.integers(10, 80)), 3
        nt = nt4_codes(synth.rand_dna(rng, int(rng.integers(30, 300))))
        pssm = rng.integers(-6, 7, size=qlen * 22).tolist()
        q = int(rng.integers(kmer - 1, qlen))
        x = int(rng.integers(3 * kmer - 1, len(nt)))
        sc, _ = pssm_ref.diag_seg(nt.tolist(), pssm, qlen, kmer, x, q, 10 ** 9, codon, ax)
        p0, q0 = x + 1 - 3 * kmer, q + 1 - kmer
        s = lambda j: pssm[j * 22 + pssm_ref._codon_aa(nt, p0 + 3 * (j - q0), codon, ax)]
        fwd = [j for j in range(q0, qlen) if 0 <= p0 + 3 * (j - q0) and p0 + 3 * (j - q0) + 2 < len(nt)]
        bwd = [j for j in range(q0 - 1, -1, -1) if 0 <= p0 + 3 * (j - q0) and p0 + 3 * (j - q0) + 2 < len(nt)]
        best = lambda js: max([0] + list(np.cumsum([s(j) for j in js])))
        assert sc == best(fwd) + best(bwd)
    for _ in range(40):
        seq = orf_case(rng, int(rng.integers(0, 100)))["seq"].tolist()
        m = int(rng.integers(1, 20))
        orfs, aa = pssm_ref.orf_walk(seq, m, codon)
        assert len(aa) == sum(k for _, k in orfs)
        found = set()
        for st, k in orfs:
            assert k >= m and all(b < 4 for b in seq[st:st + 3 * k])
            cods = [codon[seq[st + 3 * t] << 4 | seq[st + 3 * t + 1] << 2 | seq[st + 3 * t + 2]] for t in range(k)]
            assert max(cods) < 20
            for pos in (st - 3, st + 3 * k):   # the neighbouring codons of the frame end the run
                if 0 <= pos and pos + 3 <= len(seq) and all(b < 4 for b in seq[pos:pos + 3]):
                    assert codon[seq[pos] << 4 | seq[pos + 1] << 2 | seq[pos + 2]] >= 20
            found.add((st, k))
        # completeness: every maximal run of >= m good codons is reported
        for f in range(3):
            run = []
            for st in list(range(f, len(seq) - 2, 3)) + [None]:
                good = st is not None and all(b < 4 for b in seq[st:st + 3]) and \
                    codon[seq[st] << 4 | seq[st + 1] << 2 | seq[st + 2]] < 20
                if good:
                    run.append(st)
                else:
                    if len(run) >= m:
                        assert (run[0], len(run)) in found
                    run = []


def _gapped_cases():
    for c in _load("ps_gapped_cases.json.gz"):
        yield (np.frombuffer(base64.b64decode(c["pssm"]), dtype=np.int8).reshape(-1, 22),
               [np.frombuffer(base64.b64decode(o), dtype=np.uint8) for o in c["orfs"]], c)


def test_ps_gapped_golden():
    for pssm, orfs, c in _gapped_cases():
        for o, d, exp in zip(orfs, c["diags"], c["sw"]):
            assert pssm_ref.ps_gapped(pssm.ravel().tolist(), len(pssm), o.tolist(), d, c["band"], c["go"], c["ge"]) == exp


def test_ps_gapped_closed_forms():
    """a band that covers the whole matrix gives the full-matrix score; a profile that rewards
    exactly one residue per row scores its gap-free run on an ORF that spells it out; brute force
    over all local alignments on tiny inputs (no gaps cheaper than matches)"""
    from itertools import product
    rng = np.random.default_rng(5)
    for _ in range(60):
        L, n = int(rng.integers(1, 30)), int(rng.integers(1, 40))
        pssm = rng.integers(-6, 7, size=L * 22).tolist()
        orf = rng.integers(0, 20, size=n).tolist()
        full = pssm_ref.ps_gapped(pssm, L, orf, 0, 0, 11, 1)
        assert pssm_ref.ps_gapped(pssm, L, orf, int(rng.integers(-5, 5)), L + n + 5, 11, 1) == full
    for _ in range(20):
        L = int(rng.integers(1, 40))
        res = rng.integers(0, 20, size=L)
        pssm = np.full((L, 22), -4, dtype=np.int64)
        pssm[np.arange(L), res] = 5
        pre = int(rng.integers(0, 10))
        orf = rng.integers(0, 20, size=pre).tolist() + res.tolist()
        assert pssm_ref.ps_gapped(pssm.ravel().tolist(), L, orf, pre, 32, 11, 1) >= 5 * L
        assert pssm_ref.ps_gapped(pssm.ravel().tolist(), L, orf, pre, 32, 11, 1) <= 5 * L   # no better local path
    # gap-free local alignment is exact when gaps can never pay: go larger than any positive run
    for _ in range(30):
        L, n = int(rng.integers(1, 6)), int(rng.integers(1, 7))
        pssm = rng.integers(-5, 6, size=L * 22).tolist()
        orf = rng.integers(0, 20, size=n).tolist()
        brute = 0
        for i, t in product(range(L), range(n)):
            s = 0
            for d in range(min(L - i, n - t)):
                s += pssm[(i + d) * 22 + orf[t + d]]
                brute = max(brute, s)
        assert pssm_ref.ps_gapped(pssm, L, orf, 0, 0, 1000, 1000) == brute


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_ps_gapped_matches_reference_live():
    from cases import gapped_case
    ref = nsffi.Reference()
    rng = np.random.default_rng(77)
    for k in range(60):
        c = gapped_case(rng, k)
        for o, d in zip(c["orfs"], c["diags"]):
            assert pssm_ref.ps_gapped(c["pssm"].ravel().tolist(), len(c["pssm"]), o.tolist(), d, c["band"], c["go"],
                                      c["ge"]) == ref.ps_gapped(c["pssm"], o, d, c["band"], c["go"], c["ge"])


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_hmm_pssm_matches_reference():
    """synth.hmm_pssm (the PSSM the GPU search is fed) equals mp_hmm_pssm of the parsed record"""
    from protlocus_amd import synth
    ref = nsffi.Reference()
    rng = np.random.default_rng(404)
    for M in (1, 7, 60, 333):
        em, tr = synth.hmm_params(rng, M)
        em[0][3] = 0.0                                  # a '*' emission (-64)
        h = ref.hmm(synth.hmm_text(em, tr))
        try:
            assert (ref.pssm(h) == synth.hmm_pssm(em)).all()
        finally:
            ref.lib.mp_hmm_destroy(h)


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_ps_build_query_matches_one_shot_search():
    """the kept-index adaptor (one build, many queries) returns what the one-shot adaptor does"""
    from cases import prescreen_case
    ref = nsffi.Reference()
    rng = np.random.default_rng(405)
    for k in range(4):
        c = prescreen_case(rng, k)
        p = c["params"]
        h = ref.hmm(c["text"])
        hnd = ref.ps_build(c["genome"], c["ctg_off"], c["ctg_len"], kmer=p["kmer"], min_orf=p["min_orf"])
        try:
            q = {kk: v for kk, v in p.items() if kk != "min_orf"}
            want, _ = ref.ps_search(c["genome"], c["ctg_off"], c["ctg_len"], h, **p)
            assert ref.ps_query(hnd, h, **q) == want
            assert ref.ps_query(hnd, h, **q) == want          # the index is read-only across queries
        finally:
            ref.ps_free(hnd)
            ref.lib.mp_hmm_destroy(h)
