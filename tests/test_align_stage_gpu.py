"""The align stage of the map path on the device (nsh_align_stage / nsh_align_records) against the
reference's own mp_align_batch + mp_write_output, on the reduced genome-scale workload (300
proteins with planted gene models in 5 Mb; golden/config4_align_cases.json.gz).

The candidates are the reference's: each query's refined regions as its mp_map_finish hands them
to mp_align_batch (anchors, spans, extension limits). All 311 candidates of all queries go through
the stages together (one extension plan or global-DP batch per stage). Every candidate must come out
as the reference's does: dropped or aligned, the same spans, DP score and CIGAR, and the same PAF
and GFF3 records byte for byte.
"""
import base64
import ctypes as C
import gzip
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


class AlignOpt(C.Structure):
    _fields_ = [("go", C.c_int32), ("ge", C.c_int32), ("io", C.c_int32), ("io_end", C.c_int32), ("fs", C.c_int32),
                ("xdrop", C.c_int32), ("end_bonus", C.c_int32), ("max_ext", C.c_int32), ("max_intron", C.c_int32),
                ("kmer2", C.c_int32), ("end_retry_win", C.c_int32), ("ie_coef", C.c_float), ("sp_scale", C.c_float),
                ("sp_model", C.c_int32), ("sp_null_bonus", C.c_int32), ("mat", C.c_int8 * 484)]


class AlignReg(C.Structure):
    _fields_ = [("query", C.c_int32), ("vid", C.c_uint32), ("qs", C.c_int32), ("qe", C.c_int32), ("vs", C.c_int64),
                ("ve", C.c_int64), ("cnt", C.c_int32), ("chn_sc", C.c_int32), ("chn_sc_ungap", C.c_int32),
                ("a_off", C.c_int64), ("ext", C.c_uint64), ("status", C.c_int32), ("dp_score", C.c_int32),
                ("n_cigar", C.c_int32), ("cigar_off", C.c_int64), ("win_as", C.c_int64), ("win_ae", C.c_int64)]


def test_align_stage_matches_reference(product):
    from golden.make_golden import config4_reduced
    L = product.lib
    L.nsh_align_opt_default.argtypes = [C.c_void_p]
    L.nsh_align_stage.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.POINTER(C.c_uint32))]
    L.nsh_align_records.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int64, C.c_int32,
                                    C.c_char_p, C.c_int64]
    L.nsh_align_records.restype = C.c_int64
    with gzip.open(os.path.join(HERE, "golden", "config4_align_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    g, off, ln, pn, ps = config4_reduced()
    n_q = len(ps)
    aa = (C.c_char_p * n_q)(*ps)
    qn = (C.c_char_p * n_q)(*[x.encode() for x in pn])
    ql = np.array([len(x) for x in ps], np.int32)
    ctg = (C.c_char_p * len(off))(*([b"c"] * len(off)))
    regs, anchors, owner = [], [], []
    for c in cases:
        a = np.frombuffer(base64.b64decode(c["anchors"]), np.uint64)
        k = 0
        for j, r in enumerate(c["regions"]):
            vid, qs, qe, cnt, chn, chn_u, vs, ve, ext = r
            regs.append(AlignReg(query=c["q"], vid=vid, qs=qs, qe=qe, vs=vs, ve=ve, cnt=cnt, chn_sc=chn,
                                 chn_sc_ungap=chn_u, a_off=sum(len(x) for x in anchors) + k, ext=int(ext)))
            owner.append((c["q"], j))
            k += cnt
        anchors.append(a)
    anch = np.ascontiguousarray(np.concatenate(anchors), np.uint64)
    R = (AlignReg * len(regs))(*regs)
    opt = AlignOpt()
    L.nsh_align_opt_default(C.byref(opt))
    cig = C.POINTER(C.c_uint32)()
    print("align stage:", len(regs), "candidates of", n_q, "queries", flush=True)
    assert L.nsh_align_stage(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, aa, ql.ctypes.data, len(regs),
                             R, anch.ctypes.data, C.byref(opt), C.byref(cig)), product.error()
    by_q = {c["q"]: c for c in cases}
    buf = C.create_string_buffer(1 << 20)
    texts = {}
    bad = []
    for k, (q, j) in enumerate(owner):
        c = by_q[q]
        r = R[k]
        exp_st = c["status"][j]
        f = c["fields"][j]
        got_st = int(r.status)
        if got_st != exp_st:
            bad.append((q, j, "status", got_st, exp_st))
            continue
        if not got_st:
            continue
        ecig = np.frombuffer(base64.b64decode(c["cigar"]), np.uint32)
        e0 = sum(int(c["fields"][t][7]) for t in range(j) if c["status"][t])
        exp_cig = ecig[e0:e0 + int(f[7])].tolist()
        got_cig = [cig[r.cigar_off + t] for t in range(r.n_cigar)]
        got = [int(r.vid), r.qs, r.qe, r.vs, r.ve, r.dp_score]
        if got != [int(x) for x in f[:6]] or got_cig != exp_cig:
            ops = "MIDNSHP=XBFGUVE"
            cs = lambda cc: "".join("%d%s" % (x >> 4, ops[x & 0xf]) for x in cc)
            bad.append((q, j, got, f[:7], cs(got_cig), cs(exp_cig)))
            continue
        for flag, key in ((0, "paf"), (0x8, "gff")):
            n = L.nsh_align_records(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, ctg, qn, aa,
                                    ql.ctypes.data, C.byref(r), cig, C.byref(opt), flag, j, 0, buf, len(buf))
            assert n >= 0, product.error()
            texts.setdefault((q, key), []).append(buf.raw[:n])
    product.free(C.cast(cig, C.c_void_p))
    assert not bad, bad[:5]
    n_rec = 0
    for c in cases:
        for key in ("paf", "gff"):
            exp = base64.b64decode(c[key])
            got = b"".join(texts.get((c["q"], key), []))
            assert got == exp, (c["q"], key, got[:300], exp[:300])
            n_rec += exp.count(b"\n")
    assert sum(s for c in cases for s in c["status"]) >= 290 and n_rec > 1000
