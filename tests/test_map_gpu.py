"""GPU parity of the whole map path (nsh_mapper_init + nsh_map_batch: index, anchors, chains,
regions, refinement, the align stage's DPs and the records) against the reference's own program on
the reduced genome-scale workload: 300 proteins with planted gene models in 5 Mb
(synth.genome_with_genes seed 3; golden/config4_map_cases.json.gz is the stdout of the reference's
CPU build, PAF and --gff). Byte for byte, including the record ids and hit ranks."""
import base64
import ctypes as C
import gzip
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _bind(L):
    from test_align_stage_gpu import AlignOpt

    class MapOpt(C.Structure):
        _fields_ = [("dp", AlignOpt), ("max_occ", C.c_int32), ("max_gap", C.c_int32), ("bw", C.c_int32),
                    ("max_chn_max_skip", C.c_int32), ("max_chn_iter", C.c_int32), ("min_chn_cnt", C.c_int32),
                    ("min_chn_sc", C.c_int32), ("chn_coef_log", C.c_float), ("max_ava", C.c_int32),
                    ("best_n", C.c_int32), ("out_n", C.c_int32), ("mask_len", C.c_int32), ("pri_ratio", C.c_float),
                    ("mask_level", C.c_float), ("out_sim", C.c_float), ("out_cov", C.c_float), ("flag", C.c_uint32),
                    ("gff_delim", C.c_int32), ("gff_prefix", C.c_char_p), ("n_threads", C.c_int32),
                    ("prescreen", C.c_void_p), ("ps_tau", C.c_float), ("ps_audit", C.c_float), ("ps_seed", C.c_uint64),
                    ("cand_feat", C.c_void_p), ("cand_label", C.c_void_p), ("cand_cap", C.c_int64)]
    L.nsh_map_opt_default.argtypes = [C.c_void_p]
    L.nsh_mapper_init.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32]
    L.nsh_mapper_init.restype = C.c_void_p
    L.nsh_map_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_void_p]
    L.nsh_mapper_destroy.argtypes = [C.c_void_p]
    return MapOpt


def map_workload(product, cn, cs, pn, ps, flag=0, batch=None):
    """records of nsh_map_batch for proteins ps against contigs cs; batch: proteins per call"""
    L = product.lib
    MapOpt = _bind(L)
    lut = np.full(256, 4, np.uint8)
    for i, c in enumerate(b"ACGT"):
        lut[c] = i
    g = np.ascontiguousarray(lut[np.frombuffer(b"".join(cs), np.uint8)])
    ln = np.array([len(x) for x in cs], np.int64)
    names = (C.c_char_p * len(cn))(*[x.encode() for x in cn])
    m = L.nsh_mapper_init(len(cn), names, g.ctypes.data, ln.ctypes.data, 30, 6, 1, 8)
    assert m, product.error()
    o = MapOpt()
    L.nsh_map_opt_default(C.byref(o))
    o.flag = flag
    out, rid, stats = [], C.c_int64(0), np.zeros(24, np.int64)
    batch = batch or len(ps)
    try:
        for b0 in range(0, len(ps), batch):
            sub = ps[b0:b0 + batch]
            qn = (C.c_char_p * len(sub))(*[x.encode() for x in pn[b0:b0 + batch]])
            qs = (C.c_char_p * len(sub))(*sub)
            ql = np.array([len(x) for x in sub], np.int32)
            buf, n = C.c_void_p(), C.c_int64()
            st = np.zeros(24, np.int64)
            assert L.nsh_map_batch(m, len(sub), qn, qs, ql.ctypes.data, C.byref(o), C.byref(rid), C.byref(buf),
                                   C.byref(n), st.ctypes.data), product.error()
            out.append(C.string_at(buf, n.value))
            product.free(buf)
            stats += st
    finally:
        L.nsh_mapper_destroy(m)
    return b"".join(out), stats


def _first_diff(got, exp):
    gl, el = got.split(b"\n"), exp.split(b"\n")
    for k, (a, b) in enumerate(zip(gl, el)):
        if a != b:
            return k, a[:400], b[:400]
    return min(len(gl), len(el)), len(gl), len(el)


@pytest.mark.parametrize("key", ["paf", "gff"])
def test_map_path_matches_reference_program(product, key):
    from protlocus_amd import synth
    with gzip.open(os.path.join(HERE, "golden", "config4_map_cases.json.gz"), "rt") as f:
        exp = base64.b64decode(json.load(f)[key])
    cn, cs, pn, ps, _ = synth.genome_with_genes(n_prot=300, genome_len=5_000_000, n_ctg=4, seed=3)
    print("map path:", len(ps), "proteins,", key, flush=True)
    got, st = map_workload(product, cn, cs, pn, ps, flag=0x8 if key == "gff" else 0, batch=128)
    if key == "gff":
        got = b"##gff-version 3\n" + got
    assert got == exp, _first_diff(got, exp)
    assert st[5] == exp.count(b"\n" if key == "paf" else b"\tmRNA\t") and st[5] > 250, st


def test_prescreen_hook(product):
    """the learned prescreen between coarse chaining and refinement: the exported candidates carry
    finite features and one label per record; a cut that keeps everything, and an audit that keeps
    everything the cut rejects, leave the records unchanged; a cut with no audit drops them all"""
    import numpy as np
    from protlocus_amd import capi, prescreen, synth
    cn, cs, pn, ps, _ = synth.genome_with_genes(n_prot=120, genome_len=2_000_000, n_ctg=3, seed=5)
    mp = capi.Mapper(cn, cs)
    net = capi.MLP(prescreen.init_params(np.random.default_rng(0)))
    try:
        plain, st = mp.map(pn, ps)
        X, y, recs, st2 = prescreen.map_candidates(mp, pn, ps)
        assert recs == plain and st2[6] == len(y) > len(ps) and np.isfinite(X).all()
        assert int(y.sum()) == int(st[5]) > 100, (y.sum(), st[5])
        for tau, audit, same in ((-1e30, 0.0, True), (1e30, 1.0, True), (1e30, 0.0, False)):
            o = capi.MapOpt.default(prescreen=net.h, ps_tau=tau, ps_audit=audit, ps_seed=7)
            got, s = mp.map(pn, ps, o)
            print("prescreen cut", tau, audit, s[6], s[7], s[14], s[15], flush=True)
            if same:
                assert got == plain, (tau, audit)
                assert s[7] == 0 and s[15] == (st[5] if audit > 0 else 0)
            else:
                assert got == b"" and s[7] == s[6] == len(y)
    finally:
        net.close()
        mp.close()
