"""GPU parity of the k-mer block index (nsh_index_build) against the reference's mp_idx_boff +
mp_sketch_nt4 + build_bidx (index.c:14-29, 52-97; sketch.c:108-169): bucket starts, entries and
block offsets identical, on the golden records of the compiled reference and on random genomes
against the pinned restatement tests/index_ref.py."""
import base64
import gzip
import json
import os
import zlib

import numpy as np
import pytest

import index_ref
from cases import index_case, index_keys, unpack_keys

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def test_index_golden(product):
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "index_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for k, c in enumerate(cases):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        print("case", k, "genome", len(g), "contigs", len(c["ctg_len"]), "kmer", c["kmer"], flush=True)
        ki, kb, bo = capi.index_build(g, c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"],
                                      c["bbit"])
        assert bo.tolist() == c["bo"] and len(kb) == c["n_kb"], k
        assert np.array_equal(index_keys(ki, kb), unpack_keys(c["keys"])), k
        # the same genome in the reference's 4-bit form (nsh_index_build_nt4)
        pki, pkb, pbo = capi.index_build(capi.pack_nt4(g), c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"],
                                         c["mod_bit"], c["bbit"], packed=True)
        assert np.array_equal(pki, ki) and np.array_equal(pkb, kb) and np.array_equal(pbo, bo), k


def test_index_random(product):
    from protlocus_amd import capi
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    rng = np.random.default_rng(99)
    for k in range(6):
        c = index_case(rng, k, scale=4)
        print("case", k, "genome", len(c["genome"]), flush=True)
        ki, kb, bo = capi.index_build(c["genome"], c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"],
                                      c["mod_bit"], c["bbit"])
        wki, wkb, wbo, _ = index_ref.index_build(c["genome"], c["ctg_off"].tolist(), c["ctg_len"].tolist(),
                                                 c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"], t["codon"],
                                                 t["codon13"])
        assert np.array_equal(ki, wki) and kb.tolist() == list(wkb) and bo.tolist() == wbo, k


def test_index_rejects_bad_arguments(product):
    from protlocus_amd.capi import lib
    g = np.zeros(10, np.uint8)
    off, ln = np.zeros(1, np.int64), np.full(1, 10, np.int64)
    for kmer, mod_bit, bbit, mal in ((8, 1, 8, 30), (6, 24, 8, 30), (6, 1, 31, 30), (6, 1, 8, 0)):
        assert not lib().nsh_index_build(1, g.ctypes.data, off.ctypes.data, ln.ctypes.data, mal, kmer, mod_bit, bbit)


def test_anchors_golden(product):
    """nsh_index_anchors against the reference's mp_sketch_query + mp_cal_max_occ +
    mp_collect_anchors (map.c:583-597, golden records): every query's sorted anchors"""
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "anchor_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for k, c in enumerate(cases):
        g = np.frombuffer(zlib.decompress(base64.b64decode(c["genome"])), dtype=np.uint8)
        print("case", k, "genome", len(g), "queries", len(c["queries"]), "max_occ", c["max_occ"], flush=True)
        ix = capi.Index(g, c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"])
        try:
            got = ix.anchors([q.encode() for q in c["queries"]], c["max_occ"])
        finally:
            ix.close()
        for j, (a, want) in enumerate(zip(got, c["anchors"])):
            assert np.array_equal(a, unpack_keys(want)), (k, j)


def test_anchors_random(product):
    from cases import anchor_case
    from protlocus_amd import capi
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        t = json.load(f)
    rng = np.random.default_rng(123)
    for k in range(6):
        c = anchor_case(rng, k, n_query=40)
        print("case", k, "genome", len(c["genome"]), flush=True)
        ix = capi.Index(c["genome"], c["ctg_off"], c["ctg_len"], c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"])
        try:
            got = ix.anchors(c["queries"], c["max_occ"])
            assert ix.anchors([], c["max_occ"]) == []
        finally:
            ix.close()
        ki, kb, _, _ = index_ref.index_build(c["genome"], c["ctg_off"].tolist(), c["ctg_len"].tolist(),
                                             c["min_aa_len"], c["kmer"], c["mod_bit"], c["bbit"], t["codon"],
                                             t["codon13"])
        for j, q in enumerate(c["queries"]):
            want = index_ref.anchors(ki, kb, q, c["kmer"], c["mod_bit"], c["max_occ"], t["aa13"])
            assert got[j].tolist() == want, (k, j)
