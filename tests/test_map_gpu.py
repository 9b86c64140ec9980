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
                    ("gff_delim", C.c_int32), ("gff_prefix", C.c_char_p), ("n_threads", C.c_int32)]
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
    out, rid, stats = [], C.c_int64(0), np.zeros(16, np.int64)
    batch = batch or len(ps)
    try:
        for b0 in range(0, len(ps), batch):
            sub = ps[b0:b0 + batch]
            qn = (C.c_char_p * len(sub))(*[x.encode() for x in pn[b0:b0 + batch]])
            qs = (C.c_char_p * len(sub))(*sub)
            ql = np.array([len(x) for x in sub], np.int32)
            buf, n = C.c_void_p(), C.c_int64()
            st = np.zeros(16, np.int64)
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
