"""GPU parity of the --gpu-prefilter candidate selection (nsh_gapless_select_batch and the fused
nsh_gapless_prefilter_batch) against the reference's hit.c:149-263 as map.c:1075-1083 and
507-521 drive it: keep flags, K per query and the calibration must be identical. The checks
use the golden records of the compiled reference, and random mini-batches against the pinned
restatement tests/gsel_ref.py. The fused call's scores must equal ns_gpu_gapless_score_batch's."""
import gzip
import json
import os

import numpy as np
import pytest

import gsel_ref
import nsffi
from cases import gapless_select_case

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def test_gapless_select_golden(product):
    from protlocus_amd import capi
    with gzip.open(os.path.join(HERE, "golden", "gapless_select_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for k, c in enumerate(cases):
        print("case", k, "queries", len(c["k"]), "top_k", c["top_k"], "adaptive", c["adaptive"], flush=True)
        keep, ks, cal = capi.gapless_select(c["cand_off"], c["qlen"], c["scores"], c["chn"], c["vid"], c["vs"],
                                            c["top_k"], c["min_per_vid"], c["adaptive"])
        assert keep == c["keep"], k
        assert ks == c["k"], k
        assert list(cal) == c["cal"], k


def test_gapless_select_random(product):
    from protlocus_amd import capi
    rng = np.random.default_rng(2024)
    for k in range(40):
        c = gapless_select_case(rng, k)
        args = [c[x] for x in ("cand_off", "qlen", "scores", "chn", "vid", "vs")]
        got = capi.gapless_select(*args, c["top_k"], c["min_per_vid"], c["adaptive"])
        want = gsel_ref.select_batch(*[a.tolist() for a in args], c["top_k"], c["min_per_vid"], c["adaptive"])
        assert got[0] == want[0] and got[1] == want[1] and got[2] == tuple(want[2]), k


def test_gapless_select_edges(product):
    from protlocus_amd import capi
    # no queries; only empty queries; one query of 8192 candidates, every score tied
    assert capi.gapless_select([0], [], [], [], [], [], 4, 1, 1) == ([], [], (0, 0, 0.0))
    assert capi.gapless_select([0, 0, 0], [50, 60], [], [], [], [], 4, 1, 1) == ([], [0, 0], (0, 0, 0.0))
    n = 8192
    rng = np.random.default_rng(3)
    sc, chn = np.full(n, 7, np.int32), rng.integers(0, 3, n).astype(np.int32)
    vid, vs = rng.integers(0, 5, n).astype(np.uint32), rng.integers(0, 4, n).astype(np.int64)
    for top_k, mpv, ad in ((32, 2, 0), (4, 1, 1), (9, 0, 0)):
        got = capi.gapless_select([0, n], [100], sc, chn, vid, vs, top_k, mpv, ad)
        want = gsel_ref.select_batch([0, n], [100], sc.tolist(), chn.tolist(), vid.tolist(), vs.tolist(), top_k, mpv,
                                     ad)
        assert got[0] == want[0] and got[1] == want[1], (top_k, mpv, ad)
    # arguments the reference's option check refuses, and a query beyond the per-query limit
    from protlocus_amd.capi import lib
    z = np.zeros(8193, np.int64)
    assert not lib().nsh_gapless_select_batch(1, np.array([0, 1], np.int32).ctypes.data, z.ctypes.data, z.ctypes.data,
                                              z.ctypes.data, z.ctypes.data, z.ctypes.data, 2, 3, 0, z.ctypes.data,
                                              None, None)
    assert not lib().nsh_gapless_select_batch(1, np.array([0, 8193], np.int32).ctypes.data, z.ctypes.data,
                                              z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, 2, 1, 0,
                                              z.ctypes.data, None, None)


@pytest.mark.parametrize("seed,adaptive", [(1, 0), (2, 1), (3, 1)])
def test_gapless_prefilter_fused(oracle, product, seed, adaptive):
    """scores, then selection, in one call: scores equal the scoring entry point's, the selection
    equals the restatement applied to them"""
    from cases import gapless_batch, pack4
    from protlocus_amd import capi
    rng = np.random.default_rng(seed)
    codon = oracle.codon
    g, queries, cands, anchors = gapless_batch(rng, 600, codon, 5, 128)
    pk = pack4(g)
    mat = np.array(oracle.blosum62, dtype=np.int8)
    CA = (nsffi.GaplessCand * len(cands))(*cands)
    AN = (nsffi.GaplessAnchor * max(1, len(anchors)))(*anchors)
    ref_scores = np.zeros(len(cands), dtype=np.int32)
    L = product.lib
    L.ns_gpu_gapless_reset_target()
    assert L.ns_gpu_gapless_score_batch(pk, len(g), queries, len(queries), CA, len(cands), AN, max(1, len(anchors)),
                                        128, 5, 30, mat.ctypes.data, 22, codon, ref_scores.ctypes.data), product.error()
    sizes, left = [], len(cands)
    while left:
        s = min(left, int(rng.integers(1, 30)))
        sizes.append(s)
        left -= s
    off = [0] + np.cumsum(sizes).tolist()
    qlen = rng.integers(30, 600, size=len(sizes)).tolist()
    chn = rng.integers(0, 200, size=len(cands)).tolist()
    vid = rng.integers(0, 6, size=len(cands)).tolist()
    vs = rng.integers(0, 10 ** 6, size=len(cands)).tolist()
    sc, keep, ks, cal = capi.gapless_prefilter(pk, len(g), queries, CA, AN, 128, 5, 30, mat, codon, off, qlen,
                                               chn, vid, vs, 8, 1, adaptive)
    assert sc == ref_scores.tolist()
    want = gsel_ref.select_batch(off, qlen, sc, chn, vid, vs, 8, 1, adaptive)
    assert keep == want[0] and ks == want[1] and cal == tuple(want[2])
    assert 0 < sum(keep) < len(keep)
