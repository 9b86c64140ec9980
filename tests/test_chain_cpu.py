"""The anchor chaining restatement (tests/chain_ref.py, mp_chain of chain.c:7-219) pinned three
ways: the golden records of the compiled reference (synthetic anchor sets and the anchors of
real queries, with the pre-chain and main-chain parameters of map.c:611-633), fresh anchor sets
against the live reference when it is built, and closed forms: the radix sort sorts and keeps
input order for short inputs, mp_log2 stays within its approximation error of log2, the joining
score of exact and gapped anchor pairs, a perfect diagonal chains into one chain of every anchor,
and the invariants every chain list holds."""
import gzip
import json
import math
import os

import numpy as np
import pytest

import chain_ref
import nsffi
from cases import anchor_chain_case, unpack_keys

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with gzip.open(os.path.join(HERE, "golden", "anchor_chain_cases.json.gz"), "rt") as f:
        return json.load(f)


def _ints(text):
    return [int(x) for x in unpack_keys(text)]


def test_restatement_matches_golden():
    cases = _golden()
    assert len(cases) >= 40 and sum(len(_ints(c["u"])) for c in cases) > 1000
    for k, c in enumerate(cases):
        u, a = chain_ref.chain(_ints(c["anchors"]), *c["params"])
        assert u == _ints(c["u"]) and a == _ints(c["a"]), k


@pytest.mark.skipif(not nsffi.have_ref(), reason="the compiled reference (make -C oracle ref) is absent")
def test_restatement_matches_reference_live():
    ref = nsffi.Reference()
    rng = np.random.default_rng(77)
    for k in range(15):
        c = anchor_chain_case(rng, k, n_max=1200)
        assert chain_ref.chain([int(x) for x in c["anchors"]], *c["params"]) == ref.chain(c["anchors"], *c["params"]), k


def test_radix_sort_closed_forms():
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 63, 64, 65, 300, 3000):
        for bits in (8, 20, 40, 64):
            xs = [int(x) for x in rng.integers(0, 2 ** min(bits, 63), size=n, dtype=np.int64)]
            arr = [(x, i) for i, x in enumerate(xs)]
            chain_ref.radix_sort_128x(arr)
            assert [x for x, _ in arr] == sorted(xs) and sorted(i for _, i in arr) == list(range(n)), (n, bits)
            if n <= chain_ref.RS_MIN_SIZE:   # insertion sort alone: equal keys keep their input order
                assert arr == sorted(((x, i) for i, x in enumerate(xs)), key=lambda p: p[0])
    # the largest keys (a negative score widened to uint64) sort last
    arr = [(2 ** 64 - 5, 0), (7, 1), (2 ** 64 - 1, 2)] + [(3, 3 + i) for i in range(100)]
    chain_ref.radix_sort_128x(arr)
    assert [i for _, i in arr[-3:]] == [1, 0, 2]


def test_log2_closed_forms():
    for k in range(1, 30):
        assert abs(float(chain_ref.log2f(2.0 ** k)) - (k + 0.00493976)) < 1e-5, k
    for x in list(range(2, 5000)) + [10 ** 6 + 7, 123456789]:
        assert abs(float(chain_ref.log2f(x)) - math.log2(x)) < 0.0051, x


def test_comput_sc_closed_forms():
    sc = chain_ref.comput_sc
    p = dict(max_dist_x=5000, max_dist_y=5000, bw=5000, coef=0.75, is_spliced=1)
    # bbit 0: anchors on the diagonal (3 nt per residue) score min(dq, kmer)
    for d in range(1, 12):
        assert sc((100 + 3 * d) << 32 | (10 + d), 100 << 32 | 10, bbit=0, kmer=5, **p) == min(d, 5)
    # unspliced, a one-base gap costs int(0.33334 + 1), two bases int(0.66668 + 0.75 * (log2(3) - 1) + 1);
    # spliced, a target gap costs the smaller of the two terms; a residue count short of kmer
    # that is not a whole number of codons in the target loses one more
    u = dict(p, is_spliced=0)
    assert sc(109 << 32 | 13, 100 << 32 | 10, bbit=0, kmer=3, **u) == 3
    assert sc(110 << 32 | 13, 100 << 32 | 10, bbit=0, kmer=3, **u) == 3 - 1
    assert sc(111 << 32 | 13, 100 << 32 | 10, bbit=0, kmer=3, **u) == 3 - 2
    assert sc(111 << 32 | 13, 100 << 32 | 10, bbit=0, kmer=3, **p) == 3
    assert sc(110 << 32 | 13, 100 << 32 | 10, bbit=0, kmer=5, **u) == 3 - 1 - 1
    # an intron (target gap) is charged min(linear, log) when spliced, their sum otherwise
    a_i, a_j = (100 + 3 * 4 + 3000) << 32 | 14, 100 << 32 | 10
    lin, lg = 3000 * 0.33334, 0.75 * (math.log2(3001) - 1) + 1
    assert sc(a_i, a_j, bbit=0, kmer=3, **p) == 3 - int(min(lin, lg))
    assert sc(a_i, a_j, bbit=0, kmer=3, **dict(p, is_spliced=0)) == 3 - int(lin + lg)
    # a query gap is never an intron
    assert sc(112 << 32 | (14 + 1000), 100 << 32 | 10, bbit=0, kmer=3, **p) == 3 - int(
        (1004 * 3 - 12) * 0.33334 + 0.75 * (math.log2(2993) - 1) + 1)
    # bbit 8: same block and within a block's slack is a block bonus; a drop off the band is -inf
    assert sc(7 << 32 | 15, 7 << 32 | 10, bbit=8, kmer=6, **p) == 5 + chain_ref.BLOCK_BONUS
    assert sc(8 << 32 | 95, 7 << 32 | 10, bbit=8, kmer=6, **p) == 6
    assert sc(7 << 32 | 10, 7 << 32 | 10, bbit=8, kmer=6, **p) == chain_ref.INT32_MIN
    assert sc(9000 << 32 | 11, 7 << 32 | 10, bbit=8, kmer=6, **dict(p, bw=100)) == chain_ref.INT32_MIN


def test_chain_closed_forms():
    n = 200
    diag = [(3 * i) << 32 | i for i in range(n)]
    u, a = chain_ref.chain(diag, 5000, 5000, 500, 25, 5000, 2, 0, 0.75, 1, 1, 0)
    assert u == [n << 32 | n] and a == diag
    # two diagonals far apart in the target: two chains, by first target position
    far = [(10 ** 6 + 3 * i) << 32 | i for i in range(50)]
    u, a = chain_ref.chain(sorted(diag[:80] + far), 5000, 5000, 500, 25, 5000, 2, 0, 0.75, 1, 1, 0)
    assert u == [80 << 32 | 80, 50 << 32 | 50] and a == diag[:80] + far
    assert chain_ref.chain([], 1, 1, 1, 1, 1, 1, 0, 0.75, 1, 3, 8) == ([], [])
    assert chain_ref.chain(diag[:3], 5000, 5000, 500, 25, 5000, 4, 0, 0.75, 1, 1, 0) == ([], [])


def test_chain_invariants():
    rng = np.random.default_rng(9)
    for k in range(20):
        c = anchor_chain_case(rng, k, n_max=800)
        a_in = [int(x) for x in c["anchors"]]
        u, a = chain_ref.chain(a_in, *c["params"])
        min_cnt, min_sc = c["params"][5], c["params"][6]
        assert sum(x & 0xffffffff for x in u) == len(a)
        left = {}
        for x in a_in:
            left[x] = left.get(x, 0) + 1
        k0, firsts = 0, []
        for x in u:
            m, s = x & 0xffffffff, chain_ref._i32(x >> 32)
            assert m >= max(1, min_cnt) and s >= min_sc
            ch = a[k0:k0 + m]
            firsts.append(ch[0] >> 32)
            for y0, y1 in zip(ch, ch[1:]):
                assert y1 >> 32 >= y0 >> 32 and (y1 & 0xffffffff) > (y0 & 0xffffffff)
            for y in ch:
                left[y] -= 1
                assert left[y] >= 0
            k0 += m
        assert firsts == sorted(firsts), k
