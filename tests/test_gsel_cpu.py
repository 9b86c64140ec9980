"""The --gpu-prefilter selection restatement (tests/gsel_ref.py) pinned three ways: the golden
records of the compiled reference (hit.c:149-263 driven as map.c:1075-1083), fresh random
mini-batches against the live reference when it is built, and closed forms: the small-limit
streaming path equals the ranked greedy, the kept set without a per-vid floor is the top of the
ranking, and the calibration's noise is the upper median of the background scores."""
import gzip
import json
import os

import numpy as np
import pytest

import gsel_ref
import nsffi
from cases import gapless_select_case

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with gzip.open(os.path.join(HERE, "golden", "gapless_select_cases.json.gz"), "rt") as f:
        return json.load(f)


def test_restatement_matches_golden():
    for k, c in enumerate(_golden()):
        keep, ks, cal = gsel_ref.select_batch(c["cand_off"], c["qlen"], c["scores"], c["chn"], c["vid"], c["vs"],
                                              c["top_k"], c["min_per_vid"], c["adaptive"])
        assert keep == c["keep"], k
        assert ks == c["k"], k
        assert list(cal) == c["cal"], k


def test_golden_cover_the_paths():
    """the records reach calibrated noise, duplication pressure, both selection paths and the guard"""
    cs = _golden()
    assert sum(c["cal"][0] > 0 for c in cs) >= 8 and sum(c["cal"][2] > 0 for c in cs) >= 3
    big = sum(1 for c in cs for q in range(len(c["k"]))
              if min(c["cand_off"][q + 1] - c["cand_off"][q], c["k"][q]) > 8)
    assert big >= 50
    guards = 0
    for c in cs:
        for q in range(len(c["k"])):
            a, b = c["cand_off"][q], c["cand_off"][q + 1]
            if 0 < c["k"][q] < b - a and gsel_ref._chain_guard(c["scores"][a:b], c["chn"][a:b]) >= 0:
                guards += 1
    assert guards >= 50


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_restatement_matches_reference_live():
    ref = nsffi.Reference()
    rng = np.random.default_rng(77)
    for k in range(48):
        c = gapless_select_case(rng, k)
        args = [c[x] for x in ("cand_off", "qlen", "scores", "chn", "vid", "vs")]
        want = ref.gapless_select(*args, c["top_k"], c["min_per_vid"], c["adaptive"])
        got = gsel_ref.select_batch(*[a.tolist() for a in args], c["top_k"], c["min_per_vid"], c["adaptive"])
        assert (got[0], got[1], tuple(got[2])) == want, k


def _greedy(limit, min_per_vid, sc, chn, vid, vs):
    """the definition: in rank order take a candidate while fewer than min_per_vid of its vid are
    taken and fewer than limit in all, then fill with the best of the rest"""
    order = sorted(range(len(sc)), key=lambda i: gsel_ref.rank_key(i, sc, chn, vid, vs))
    taken, per = [], {}
    if min_per_vid > 0:
        for i in order:
            if len(taken) < limit and per.get(vid[i], 0) < min_per_vid:
                taken.append(i)
                per[vid[i]] = per.get(vid[i], 0) + 1
    for i in order:
        if len(taken) < limit and i not in taken:
            taken.append(i)
    return taken


def test_selection_closed_forms():
    rng = np.random.default_rng(5)
    for t in range(3000):
        n = int(rng.integers(1, 14))
        sc = rng.integers(0, 6, size=n).tolist()
        chn = rng.integers(0, 4, size=n).tolist()
        vid = rng.integers(0, 3, size=n).tolist()
        vs = rng.integers(0, 3, size=n).tolist()
        top_k = int(rng.integers(1, 12))
        mpv = int(rng.integers(0, 4))
        limit = min(n, top_k)
        got = gsel_ref.select_gapless(top_k, mpv, sc, chn, vid, vs)
        guard = gsel_ref._chain_guard(sc, chn) if limit < n else -1
        want = set(_greedy(limit, mpv, sc, chn, vid, vs)) | ({guard} if guard >= 0 else set())
        assert set(got) == want, (t, n, top_k, mpv)
        if mpv == 0:   # no floor: the top `limit` of the ranking, plus the guard
            order = sorted(range(n), key=lambda i: gsel_ref.rank_key(i, sc, chn, vid, vs))
            assert set(got) - {guard} <= set(order[:limit]) and len(got) in (limit, limit + 1)
    # the guard: the best chain survives a gapless ranking that would drop it
    assert gsel_ref.select_gapless(1, 0, [10, 9], [50, 100], [0, 0], [0, 1]) == [0, 1]
    assert gsel_ref.select_gapless(1, 0, [10, 9], [80, 100], [0, 0], [0, 1]) == [0]


def test_calibration_closed_forms():
    rng = np.random.default_rng(6)
    for t in range(200):
        n_q = int(rng.integers(1, 60))
        sizes = rng.integers(0, 8, size=n_q)
        off = [0] + np.cumsum(sizes).tolist()
        sc = rng.integers(-3, 40, size=off[-1]).tolist()
        noise, n_scored, dup = gsel_ref.calibrate(off, sc)
        bg, scored, conf, multi = [], 0, 0, 0
        for q in range(n_q):
            s = sorted(sc[off[q]:off[q + 1]], reverse=True)
            if len(s) < 2 or s[0] <= 0:
                continue
            scored += 1
            bg += [x for x in s if 0 < x < s[0]]
        med = int(np.sort(bg)[len(bg) // 2]) if len(bg) >= 64 else 0
        for q in range(n_q):
            s = sorted(sc[off[q]:off[q + 1]], reverse=True)
            if len(s) < 2 or s[0] <= 0 or s[0] < 4 * med:
                continue
            conf += 1
            multi += s[1] > 0 and s[1] * 100 >= s[0] * 80
        assert (noise, n_scored) == (med, scored), t
        assert dup == (float(np.float32(multi) / np.float32(conf)) if conf else 0.0), t


def test_planted_mistakes_fail():
    """a comparator with vid descending changes the golden outcome; a tied winner counts as its
    own runner-up (hit.c:159-162), so one query of two equal scores is a near tie"""
    cs = _golden()
    orig = gsel_ref.rank_key
    try:
        gsel_ref.rank_key = lambda i, sc, chn, vid, vs: (-sc[i], -chn[i], -vid[i], vs[i], i)
        assert any(gsel_ref.select_batch(c["cand_off"], c["qlen"], c["scores"], c["chn"], c["vid"], c["vs"],
                                         c["top_k"], c["min_per_vid"], c["adaptive"])[0] != c["keep"] for c in cs)
    finally:
        gsel_ref.rank_key = orig
    assert gsel_ref.calibrate([0, 2], [7, 7]) == (0, 1, 1.0)
    assert gsel_ref.calibrate([0, 2], [7, 5]) == (0, 1, 0.0)
