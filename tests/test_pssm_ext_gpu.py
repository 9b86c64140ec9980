"""GPU parity of the extension DP scored with a per-request PSSM (HMM mode: align.c:378-380 sets
ns_opt_ext.pssm to the query profile's row of the first extended residue, mp_pssm_at).

The PSSM rows are indexed in query order and read reversed for a left extension (ns_pssm_sc,
nasw-sse.c:217-221). Checked, bit-exact (nt_len, aa_len, score):
  * every golden PSSM case the reference's nasw-sse.c wrote (golden/dp_cases.json.gz: CIGAR,
    score-only, left and right extensions, int16 and int32);
  * single ns_global_gs16b calls with random profiles, left and right, protein lengths across the
    column-count groups of the extension kernel (C 1-4, 5-8, 9-16), against the oracle;
  * one ns_ext_batch_* flush whose requests carry different PSSMs (and some none): they share one
    device plan, and every result equals the oracle's;
  * the mp_align_step pattern: one query profile, a left extension from the query start and a right
    extension from row qe (pointer into the same profile), through the batch API.
"""
import base64
import ctypes as C
import gzip
import json
import os

import numpy as np
import pytest

import nsffi
from cases import planted_case
from nsffi import NS_F_EXT_LEFT, NS_F_EXT_RIGHT

pytestmark = pytest.mark.gpu


def _ora(oracle, product, ns, a, kw):
    o = oracle.opt(**kw)
    w = product.lib.ns_ext_stripe_width(C.byref(product.opt(**kw)), len(a))
    return oracle.ext(ns, a, o, None, (len(a) + w - 1) // w * w)[0]


def test_golden_pssm_cases(product):
    with gzip.open(os.path.join(os.path.dirname(__file__), "golden", "dp_cases.json.gz"), "rt") as f:
        cases = [c for c in json.load(f) if "pssm" in c["kw"]]
    assert len(cases) >= 16
    modes = set()
    for c in cases:
        kw = dict(c["kw"])
        kw["pssm"] = np.array(kw["pssm"], dtype=np.int8)
        ss = None if c["ss"] is None else base64.b64decode(c["ss"])
        got = product.run(c["fn"], c["ns"].encode(), c["as"], product.opt(**kw), ss)
        assert list(got) == c["out"], (c["fn"], kw["flag"], len(c["as"]))
        modes.add((c["fn"], kw["flag"]))
    assert (("ns_global_gs16b", NS_F_EXT_LEFT) in modes) and (("ns_global_gs16b", NS_F_EXT_RIGHT) in modes)


def test_pssm_extension_single_calls(oracle, product):
    rng = np.random.default_rng(4242)
    for al in (1, 9, 16, 17, 100, 255, 300, 511, 600, 1023, 1100, 1600):
        ns, a = planted_case(rng, al)
        pssm = rng.integers(-7, 10, size=(al, 22)).astype(np.int8)
        for flag in (NS_F_EXT_LEFT, NS_F_EXT_RIGHT):
            kw = dict(flag=flag, pssm=pssm)
            got = product.run("ns_global_gs16b", ns, a, product.opt(**kw))[:3]
            assert got == _ora(oracle, product, ns, a, kw), (al, flag)


def test_pssm_batch_one_plan(oracle, product):
    L = product.lib
    rng = np.random.default_rng(77)
    b = L.ns_ext_batch_init(None)
    assert b, product.error()
    reqs, keep = [], []
    for k in range(40):
        al = int(rng.integers(5, 700))
        ns, a = planted_case(rng, al)
        flag = NS_F_EXT_LEFT if k & 1 else NS_F_EXT_RIGHT
        kw = dict(flag=flag)
        if k % 5:                       # every fifth request scores with the matrix
            kw["pssm"] = rng.integers(-6, 9, size=(al, 22)).astype(np.int8)
        if k % 3 == 0:
            kw["io"] = 19                 # io is per request as well
        o = product.opt(**kw)
        ab = a.encode()                   # the batch borrows the strings until the result is fetched
        keep.append((o, ab))
        reqs.append((L.ns_ext_batch_add(b, ns, len(ns), ab, len(a), C.byref(o), None), ns, a, kw))
    before = L.ns_ext_gpu_launch_count()
    assert L.ns_ext_batch_flush(b) == len(reqs)
    for rid, ns, a, kw in reqs:
        r = L.ns_ext_batch_get(b, rid).contents
        assert (r.nt_len, r.aa_len, r.score) == _ora(oracle, product, ns, a, kw), (len(a), kw["flag"], "pssm" in kw)
    assert L.ns_ext_gpu_launch_count() - before == 1     # the PSSM pointer does not split the plan
    L.ns_ext_batch_destroy(b)


def test_pssm_profile_rows_like_align_step(oracle, product):
    """left extension over aa[0, as1) with the profile's row 0, right extension over aa[qe, len)
    with the row of qe: both pointers into one query profile"""
    L = product.lib
    rng = np.random.default_rng(5)
    b = L.ns_ext_batch_init(None)
    reqs, keep = [], []
    for k in range(12):
        qlen = int(rng.integers(60, 500))
        prof = rng.integers(-6, 9, size=(qlen, 22)).astype(np.int8)
        ns, a = planted_case(rng, qlen)
        as1 = int(rng.integers(1, qlen // 2))
        qe = int(rng.integers(qlen // 2, qlen - 1))
        for flag, sub, rows in ((NS_F_EXT_LEFT, a[:as1], prof[:as1]), (NS_F_EXT_RIGHT, a[qe:], prof[qe:])):
            o = product.opt(flag=flag)
            o.pssm = prof.ctypes.data + (0 if flag == NS_F_EXT_LEFT else qe * 22)
            sb = sub.encode()             # borrowed by the batch until the result is fetched
            keep.append((o, prof, sb))
            reqs.append((L.ns_ext_batch_add(b, ns, len(ns), sb, len(sub), C.byref(o), None), ns, sub,
                         dict(flag=flag, pssm=np.ascontiguousarray(rows))))
    assert L.ns_ext_batch_flush(b) == len(reqs)
    for rid, ns, sub, kw in reqs:
        r = L.ns_ext_batch_get(b, rid).contents
        assert (r.nt_len, r.aa_len, r.score) == _ora(oracle, product, ns, sub, kw)
    L.ns_ext_batch_destroy(b)
