"""The k-mer block index and anchor restatements (tests/index_ref.py) pinned three ways: the golden
records of the compiled reference (index.c:14-29, 52-97; sketch.c:7-38, 108-169; map.c:304-395,
583-597), fresh genomes and queries against the live reference when it is built, and closed forms:
the hash is a bijection inside its mask, a hand-built ORF yields exactly its k-mers, every
bucket's entries rise, and a query made of an indexed k-mer finds that k-mer's blocks."""
import base64
import gzip
import json
import os
import zlib

import numpy as np
import pytest

import index_ref
import nsffi
from cases import anchor_case, index_case, index_keys, unpack_keys

HERE = os.path.dirname(os.path.abspath(__file__))


def _codon13():
    """ns_tab_codon and ns_tab_codon13 of the golden tables"""
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    return t["codon"], t["codon13"]


def _golden():
    with gzip.open(os.path.join(HERE, "golden", "index_cases.json.gz"), "rt") as f:
        return json.load(f)


def test_restatement_matches_golden():
    codon, codon13 = _codon13()
    for k, c in enumerate(_golden()):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        ki, kb, bo, keys = index_ref.index_build(g, c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"],
                                                 c["mod_bit"], c["bbit"], codon, codon13)
        assert bo == c["bo"] and len(kb) == c["n_kb"], k
        assert np.array_equal(np.array(keys, dtype=np.uint64), unpack_keys(c["keys"])), k
        assert np.array_equal(index_keys(ki, kb), unpack_keys(c["keys"])), k


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_restatement_matches_reference_live():
    ref = nsffi.Reference()
    import ctypes as C
    codon = list((C.c_uint8 * 64).in_dll(ref.lib, "ns_tab_codon"))
    codon13 = list((C.c_uint8 * 64).in_dll(ref.lib, "ns_tab_codon13"))
    rng = np.random.default_rng(31)
    for k in range(6):
        c = index_case(rng, k)
        want = ref.index(c["genome"], c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"])
        ki, kb, bo, _ = index_ref.index_build(c["genome"], c["ctg_off"].tolist(), c["ctg_len"].tolist(),
                                              c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"], codon, codon13)
        assert np.array_equal(ki, want[0]) and list(kb) == want[1].tolist() and bo == want[2].tolist(), k


def test_closed_forms():
    codon, codon13 = _codon13()
    # the hash permutes the keys inside its mask
    for bits in (4, 8, 12):
        mask = (1 << bits) - 1
        assert sorted(index_ref.hash32_mask(x, mask) for x in range(mask + 1)) == list(range(mask + 1))
    # GCT repeated 40 times between two Ns: each frame is one stop-free run (GCT x 40, CTG x 39,
    # TGC x 39) whose k-mers are all alike: one entry per (frame, block)
    seq = [4] + [2, 1, 3] * 40 + [4]
    kmer, bbit = 3, 4
    want = set()
    for f in (1, 2, 3):
        n = (121 - f) // 3
        a = codon13[seq[f] << 4 | seq[f + 1] << 2 | seq[f + 2]]
        y = index_ref.hash32_mask((a << 8) | (a << 4) | a, 0xfff)
        want |= {(y << 32) | ((f + 3 * t + 2) >> bbit) for t in range(kmer - 1, n)}
    assert index_ref.sketch_nt4(seq, 30, kmer, 0, bbit, 0, codon, codon13) == sorted(want)
    assert index_ref.sketch_nt4(seq, 41, kmer, 0, bbit, 0, codon, codon13) == []   # the ORF is too short
    # an N splits every frame; no half reaches 30 codons
    seq2 = [2, 1, 3] * 25 + [4] + [2, 1, 3] * 25
    assert index_ref.sketch_nt4(seq2, 30, kmer, 0, bbit, 0, codon, codon13) == []
    # a stop (TAA) splits only the frame it lies in: the other two frames keep their long runs
    seq3 = [2, 1, 3] * 25 + [3, 0, 0] + [2, 1, 3] * 25
    frames = {(k & 0xffffffff) for k in index_ref.sketch_nt4(seq3, 30, kmer, 0, 0, 0, codon, codon13)}
    assert frames and all((e - 2) % 3 != 0 for e in frames)
    # every bucket's entries rise (the strands go in order and their blocks do too)
    for c in _golden()[:4]:
        keys = unpack_keys(c["keys"])
        assert np.all(keys[1:] > keys[:-1])


def _anchor_golden():
    with gzip.open(os.path.join(HERE, "golden", "anchor_cases.json.gz"), "rt") as f:
        return json.load(f)


def test_anchor_restatement_matches_golden():
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    for k, c in enumerate(_anchor_golden()):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        ki, kb, _, _ = index_ref.index_build(g, c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"],
                                             c["bbit"], t["codon"], t["codon13"])
        for j, (q, want) in enumerate(zip(c["queries"], c["anchors"])):
            got = index_ref.anchors(ki, kb, q.encode(), c["kmer"], c["mod_bit"], c["max_occ"], t["aa13"])
            assert np.array_equal(np.array(got, dtype=np.uint64), unpack_keys(want)), (k, j)


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_anchor_restatement_matches_reference_live():
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    ref = nsffi.Reference()
    rng = np.random.default_rng(41)
    for k in range(4):
        c = anchor_case(rng, k)
        want = ref.anchors(c["genome"], c["ctg_off"], c["ctg_len"], c["queries"], c["min_aa_len"], c["kmer"],
                           c["mod_bit"], c["bbit"], c["max_occ"])
        ki, kb, _, _ = index_ref.index_build(c["genome"], c["ctg_off"].tolist(), c["ctg_len"].tolist(),
                                             c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"], t["codon"],
                                             t["codon13"])
        for q, w in zip(c["queries"], want):
            assert index_ref.anchors(ki, kb, q, c["kmer"], c["mod_bit"], c["max_occ"], t["aa13"]) == [int(x) for x in w]


def test_anchor_closed_forms():
    """a query equal to an ORF's translation finds, at its last position, the blocks of its
    final k-mer; residues outside the 13-letter alphabet break the k-mers around them"""
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    from protlocus_amd import synth
    rng = np.random.default_rng(8)
    g = np.array([2, 1, 3] * 60, dtype=np.uint8)   # GCT x 60: an alanine run in frame 0
    ki, kb, bo, keys = index_ref.index_build(g, [0], [len(g)], 30, 3, 0, 4, t["codon"], t["codon13"])
    q = b"A" * 10
    a = index_ref.anchors(ki, kb, q, 3, 0, 20000, t["aa13"])
    y = index_ref.hash32_mask((t["codon13"][2 << 4 | 1 << 2 | 3] * 0x111) & 0xfff, 0xfff)
    blocks = [int(kb[j]) for j in range(ki[y], ki[y + 1])]
    assert blocks and a == sorted(b << 32 | p for p in range(2, 10) for b in blocks)
    assert index_ref.sketch_prot(b"AA*AA", 3, 0, t["aa13"]) == []
    assert index_ref.sketch_prot(b"", 3, 0, t["aa13"]) == []
    assert len(index_ref.sketch_prot(synth.rand_protein(rng, 50), 3, 0, t["aa13"])) == 48


def test_pack_nt4_layout():
    """capi.pack_nt4 stores base i as mp_ntseq_read does (ntseq.c:43: seq[i >> 1] |= b << (i & 1) * 4)"""
    from protlocus_amd import capi
    rng = np.random.default_rng(4)
    for n in (0, 1, 2, 7, 1000):
        codes = rng.integers(0, 5, n).astype(np.uint8)
        pk = capi.pack_nt4(codes)
        assert len(pk) == (n + 1) // 2
        assert [(int(pk[i >> 1]) >> ((i & 1) * 4)) & 15 for i in range(n)] == codes.tolist()
