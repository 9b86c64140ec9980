"""ctypes bindings shared by the tests and bench.py.

Three libraries expose the reference's DP interface (types as nasw.h:62-82):
  * the product, protlocus_amd/libnasw_hip.so (GPU kernels behind the ns_* C-ABI);
  * the oracle, oracle/liboracle.so (or_* restatement; checker only);
  * the compiled reference, oracle/_ref/libnsref.so (built from the reference sources in
    the build container when they are present; used to pin the oracle).
"""
import ctypes as C
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libnsref.so")
PRODUCT_SO = os.path.join(ROOT, "protlocus_amd", "libnasw_hip.so")

NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT, NS_F_GPU = 1, 2, 4, 8
NS_S_NONE, NS_S_GENERIC, NS_S_MAMMAL = 0, 1, 2
CIGAR_STR = "MIDNSHP=XBFGUVE"


class NsOpt(C.Structure):
    _fields_ = [("flag", C.c_int32), ("go", C.c_int32), ("ge", C.c_int32), ("io", C.c_int32), ("fs", C.c_int32),
                ("xdrop", C.c_int32), ("end_bonus", C.c_int32), ("asize", C.c_int32), ("gpu_device", C.c_int32),
                ("gpu_min_cells", C.c_int64), ("sp", C.c_int32 * 6), ("sp_null_bonus", C.c_int32),
                ("ie_coef", C.c_float), ("sc", C.c_void_p), ("pssm", C.c_void_p),
                ("nt4", C.c_void_p), ("aa20", C.c_void_p), ("codon", C.c_void_p)]


class NsRst(C.Structure):
    _fields_ = [("n_cigar", C.c_int32), ("m_cigar", C.c_int32), ("nt_len", C.c_int32), ("aa_len", C.c_int32),
                ("score", C.c_int32), ("cigar", C.POINTER(C.c_uint32))]


class OrExt(C.Structure):
    _fields_ = [("nt_len", C.c_int32), ("aa_len", C.c_int32), ("score", C.c_int32)]


class GaplessCand(C.Structure):
    _fields_ = [("target_offset", C.c_int64), ("target_length", C.c_int64), ("query_offset", C.c_int64),
                ("query_length", C.c_int32), ("anchor_offset", C.c_int32), ("anchor_count", C.c_int32),
                ("reverse", C.c_int32)]


class GaplessAnchor(C.Structure):
    _fields_ = [("target_block", C.c_int64), ("query_end", C.c_int32), ("reserved", C.c_int32)]


class GaplessProfile(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("candidates", C.c_uint64), ("anchors", C.c_uint64),
                ("seed_positions", C.c_uint64)]


class OrHmm(C.Structure):
    _fields_ = [("len", C.c_int32), ("emit", C.POINTER(C.c_int16)), ("trans", C.POINTER(C.c_int16)),
                ("pssm", C.POINTER(C.c_int8))]


OPT_FIELDS = ("flag", "go", "ge", "io", "fs", "xdrop", "end_bonus", "asize", "gpu_min_cells", "sp_null_bonus", "ie_coef")


def cigar_str(rst):
    return "".join("%d%s" % (rst.cigar[k] >> 4, CIGAR_STR[rst.cigar[k] & 0xf]) for k in range(rst.n_cigar))


def _buf(x):
    """bytes-like -> (ctypes char pointer, keepalive)"""
    if x is None:
        return None, None
    if isinstance(x, str):
        x = x.encode()
    b = C.create_string_buffer(bytes(x), len(x) + 1)
    return b, b


class _Lib:
    def opt(self, **kw):
        o = NsOpt()
        self._opt_init(C.byref(o))
        if "sp_model" in kw:
            self._opt_set_sp(C.byref(o), kw.pop("sp_model"))
        if "sp" in kw:
            for i, v in enumerate(kw.pop("sp")):
                o.sp[i] = v
        mat = kw.pop("mat", None)
        pssm = kw.pop("pssm", None)
        for k, v in kw.items():
            setattr(o, k, v)
        o._keep = []
        if mat is not None:
            m = (C.c_int8 * len(mat))(*[int(v) for v in mat])
            o.sc = C.cast(m, C.c_void_p).value
            o._keep.append(m)
        if pssm is not None:
            p = np.ascontiguousarray(pssm, dtype=np.int8)
            o.pssm = p.ctypes.data
            o._keep.append(p)
        return o


class Oracle(_Lib):
    def __init__(self, path=ORACLE_SO):
        self.lib = L = C.CDLL(path)
        L.or_make_tables.argtypes = [C.c_int]
        L.or_opt_init.argtypes = [C.c_void_p]
        L.or_opt_set_sp.argtypes = [C.c_void_p, C.c_int32]
        L.or_dp16_score.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p,
                                    C.c_int32, C.c_void_p, C.c_void_p]
        for f in (L.or_global16,):
            f.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_int32, C.c_void_p]
        for f in (L.or_global32,):
            f.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_void_p]
        L.or_global_auto.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_int32, C.c_void_p]
        L.or_global_trace.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_int32,
                                      C.c_void_p, C.POINTER(C.c_int32)]
        L.or_global_trace.restype = C.c_int32
        L.or_ie_penalty.argtypes = [C.c_float, C.c_int32]
        L.or_ie_penalty.restype = C.c_int32
        L.or_hmm_parse.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p]
        L.or_hmm_viterbi.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.or_hmm_viterbi.restype = C.c_int32
        L.or_hmm_free.argtypes = [C.c_void_p]
        L.or_gapless_score.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_void_p, C.c_int32, C.c_void_p]
        L.or_gapless_score.restype = C.c_int32
        L.or_pssm_ungapped.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.or_pssm_ungapped.restype = C.c_int32
        assert L.or_make_tables(1) == 0
        self._opt_init = L.or_opt_init
        self._opt_set_sp = L.or_opt_set_sp
        self.nt4 = bytes((C.c_uint8 * 256).in_dll(L, "or_nt4"))
        self.aa20 = bytes((C.c_uint8 * 256).in_dll(L, "or_aa20"))
        self.codon = bytes((C.c_uint8 * 64).in_dll(L, "or_codon"))
        self.blosum62 = list((C.c_int8 * 484).in_dll(L, "or_blosum62"))

    def ext(self, ns, as_, opt, ss=None, npad=None):
        """16-bit score-only DP (extension when opt.flag has EXT_*); returns (nt_len, aa_len, score)."""
        al = len(as_)
        if npad is None:
            npad = (al + 15) // 16 * 16 if al >= 16 else (al + 7) // 8 * 8
        e = OrExt()
        rows = C.c_int32()
        b_ss, _ = _buf(ss)
        self.lib.or_dp16_score(bytes(ns), len(ns), bytes(as_.encode() if isinstance(as_, str) else as_), al,
                               C.byref(opt), b_ss, npad, C.byref(e), C.byref(rows))
        return (e.nt_len, e.aa_len, e.score), rows.value

    def global_dp(self, ns, as_, opt, ss=None, mode="auto", ext_width=16):
        r = NsRst()
        a = as_.encode() if isinstance(as_, str) else bytes(as_)
        b_ss, _ = _buf(ss)
        if mode == "auto":
            self.lib.or_global_auto(bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, ext_width, C.byref(r))
        elif mode == 16:
            self.lib.or_global16(bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, ext_width, C.byref(r))
        else:
            self.lib.or_global32(bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, C.byref(r))
        out = (r.nt_len, r.aa_len, r.score, cigar_str(r))
        return out

    def trace(self, ns, as_, opt, is32=0, ss=None):
        a = as_.encode() if isinstance(as_, str) else bytes(as_)
        vs = 4 if is32 else 8
        npad = (len(a) + vs - 1) // vs * vs
        buf = np.zeros(len(ns) * npad, dtype=np.uint16)
        sc = C.c_int32()
        b_ss, _ = _buf(ss)
        self.lib.or_global_trace(bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, is32, buf.ctypes.data, C.byref(sc))
        return buf.reshape(len(ns), npad), sc.value

    def ie_penalty(self, coef, d):
        return self.lib.or_ie_penalty(coef, d)


class Reference(_Lib):
    """the reference's own nasw-sse.c/nasw-tab.c/hmm.c compiled by oracle/Makefile"""

    def __init__(self, path=REF_SO):
        self.lib = L = C.CDLL(path)
        L.ns_make_tables.argtypes = [C.c_int]
        L.ns_opt_init.argtypes = [C.c_void_p]
        L.ns_opt_set_sp.argtypes = [C.c_void_p, C.c_int32]
        for f in (L.ns_global_gs16b, L.ns_global_gs32b, L.ns_global_gs_auto):
            f.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_void_p]
        L.ns_ext_stripe_width.argtypes = [C.c_void_p, C.c_int32]
        L.ns_ext_ie_penalty.argtypes = [C.c_void_p, C.c_int32]
        L.mp_hmm_open.argtypes = [C.c_char_p]
        L.mp_hmm_open.restype = C.c_void_p
        L.mp_hmm_read1.argtypes = [C.c_void_p]
        L.mp_hmm_read1.restype = C.c_void_p
        L.mp_hmm_close.argtypes = [C.c_void_p]
        L.mp_hmm_destroy.argtypes = [C.c_void_p]
        L.mp_hmm_dp_score.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.mp_hmm_dp_score.restype = C.c_int32
        L.mp_hmm_length.argtypes = [C.c_void_p]
        L.mp_hmm_emissions.argtypes = [C.c_void_p]
        L.mp_hmm_emissions.restype = C.POINTER(C.c_int16)
        L.mp_hmm_transitions.argtypes = [C.c_void_p]
        L.mp_hmm_transitions.restype = C.POINTER(C.c_int16)
        L.mp_hmm_pssm.argtypes = [C.c_void_p]
        L.mp_hmm_pssm.restype = C.POINTER(C.c_int8)
        L.or_ref_output.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                    C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.c_char_p, C.c_int64,
                                    C.c_int32, C.c_int32, C.c_char_p, C.c_int64]
        L.or_ref_output.restype = C.c_int64
        L.or_ref_diag_seg.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.c_int32,
                                      C.c_int32, C.c_void_p]
        L.or_ref_chain_pssm_ev.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_void_p, C.c_int32, C.c_int32]
        L.or_ref_orfs.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                  C.c_int64, C.c_void_p, C.c_void_p]
        L.or_ref_orfs.restype = C.c_int64
        L.or_ref_ps_gapped.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32]
        L.or_ref_ps_search.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                       C.c_float, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_ref_ps_kmer_thr.argtypes = [C.c_int32, C.c_float, C.c_int32]
        L.or_ref_ps_build.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        L.or_ref_ps_build.restype = C.c_void_p
        L.or_ref_ps_query.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_void_p, C.c_void_p]
        L.or_ref_ps_free.argtypes = [C.c_void_p]
        L.or_ref_gapless_select.argtypes = [C.c_int32] + [C.c_void_p] * 6 + [C.c_int32] * 3 + [C.c_void_p] * 3
        L.or_ref_index_build.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4
        L.or_ref_index_build.restype = C.c_void_p
        L.or_ref_index_get.argtypes = [C.c_void_p] * 4
        L.or_ref_index_get.restype = C.c_int64
        L.or_ref_index_free.argtypes = [C.c_void_p]
        L.or_ref_anchors.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int64, C.c_void_p]
        L.or_ref_anchors.restype = C.c_int64
        L.or_ref_chain.argtypes = [C.c_int32] * 7 + [C.c_float, C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_void_p,
                                                     C.c_int32, C.c_void_p, C.c_int64, C.c_void_p]
        L.or_ref_chain.restype = C.c_int32
        assert L.ns_make_tables(1) == 0
        self._opt_init = L.ns_opt_init
        self._opt_set_sp = L.ns_opt_set_sp
        self.free = C.CDLL(None).free
        self.free.argtypes = [C.c_void_p]

    def run(self, fn, ns, as_, opt, ss=None):
        r = NsRst()
        a = as_.encode() if isinstance(as_, str) else bytes(as_)
        b_ss, _ = _buf(ss)
        getattr(self.lib, fn)(None, bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, C.byref(r))
        out = (r.nt_len, r.aa_len, r.score, cigar_str(r))
        if r.cigar:
            self.free(C.cast(r.cigar, C.c_void_p))
        return out

    def stripe_width(self, opt, al):
        return self.lib.ns_ext_stripe_width(C.byref(opt), al)

    def diag_seg(self, nt, pssm, qlen, kmer, x, q, xdrop):
        nt = np.ascontiguousarray(nt, dtype=np.uint8)
        ps = np.ascontiguousarray(pssm, dtype=np.int8)
        e = C.c_int32()
        sc = self.lib.or_ref_diag_seg(nt.ctypes.data, len(nt), ps.ctypes.data, qlen, kmer, x, q, xdrop, C.byref(e))
        return sc, e.value

    def chain_pssm_ev(self, nt, pssm, qlen, kmer, anchors, min_seg, xdrop):
        nt = np.ascontiguousarray(nt, dtype=np.uint8)
        ps = np.ascontiguousarray(pssm, dtype=np.int8)
        a = np.ascontiguousarray(anchors, dtype=np.uint64)
        return self.lib.or_ref_chain_pssm_ev(nt.ctypes.data, len(nt), ps.ctypes.data, qlen, kmer, len(a),
                                             a.ctypes.data, min_seg, xdrop)

    def ps_gapped(self, pssm, orf, diag, band, go, ge):
        ps = np.ascontiguousarray(pssm, dtype=np.int8)
        o = np.ascontiguousarray(orf, dtype=np.uint8)
        return self.lib.or_ref_ps_gapped(ps.ctypes.data, len(ps) // 22 if ps.ndim == 1 else ps.shape[0],
                                         o.ctypes.data, len(o), diag, band, go, ge)

    def hmm(self, text):
        """an mp_hmm_t parsed by the reference's mp_hmm_read1 from HMMER3 text"""
        import tempfile
        with tempfile.NamedTemporaryFile("w", suffix=".hmm", delete=False) as f:
            f.write(text)
        fp = self.lib.mp_hmm_open(f.name.encode())
        h = self.lib.mp_hmm_read1(fp)
        self.lib.mp_hmm_close(fp)
        os.unlink(f.name)
        return h

    def pssm(self, h):
        n = self.lib.mp_hmm_length(h)
        return np.ctypeslib.as_array(self.lib.mp_hmm_pssm(h), shape=(n * 22,)).reshape(n, 22).copy()

    def ps_search(self, genome, ctg_off, ctg_len, h, kmer=6, min_orf=30, sens=7.5, kmer_thr=0, kmer_max=65536,
                  min_ungapped=15, max_seqs=4000, band=32, go=11, ge=1, max_evalue=0.0):
        """the reference's mp_prescreen_build + ps_query_worker: candidate tuples, (ORFs, residues)"""
        from protlocus_amd.capi import PsCand
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        cap = 200000
        out = (PsCand * cap)()
        n_orf, n_aa = C.c_int64(), C.c_int64()
        n = self.lib.or_ref_ps_search(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, h, kmer, min_orf, sens,
                                      kmer_thr, kmer_max, min_ungapped, max_seqs, band, go, ge, max_evalue, cap, out,
                                      C.byref(n_orf), C.byref(n_aa))
        assert n <= cap
        return [out[i].tuple() for i in range(n)], (n_orf.value, n_aa.value)

    def ps_build(self, genome, ctg_off, ctg_len, kmer=6, min_orf=30):
        """the reference's mp_prescreen_build over one genome, kept for ps_query; free with ps_free"""
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        return self.lib.or_ref_ps_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, kmer, min_orf)

    def ps_query(self, handle, h, kmer=6, sens=7.5, kmer_max=65536, min_ungapped=15, max_seqs=4000, band=32, go=11,
                 ge=1, max_evalue=0.0, cap=65536, stats=None):
        """the reference's ps_query_worker for one profile over a ps_build index: candidate tuples;
        stats (a list) receives the worker's 7 counters when given"""
        from protlocus_amd.capi import PsCand
        out = (PsCand * cap)()
        st = (C.c_int64 * 7)()
        n = self.lib.or_ref_ps_query(handle, h, kmer, sens, kmer_max, min_ungapped, max_seqs, band, go, ge, max_evalue,
                                     cap, out, st)
        assert n <= cap
        if stats is not None:
            stats[:] = list(st)
        return [out[i].tuple() for i in range(n)]

    def ps_free(self, handle):
        self.lib.or_ref_ps_free(handle)

    def gapless_select(self, cand_off, qlen, scores, chn, vid, vs, top_k, min_per_vid, adaptive):
        """the reference's --gpu-prefilter selection (oracle/hit_ref.c): keep flags, K per query,
        (noise, n_query, dup_frac)"""
        off = np.ascontiguousarray(cand_off, dtype=np.int32)
        arrs = [np.ascontiguousarray(x, dtype=t) for x, t in ((qlen, np.int32), (scores, np.int32), (chn, np.int32),
                                                               (vid, np.uint32), (vs, np.int64))]
        n_q, n = len(off) - 1, int(off[-1])
        keep, ks = np.zeros(max(1, n), np.uint8), np.zeros(max(1, n_q), np.int32)
        cal = (C.c_int32 * 3)()
        self.lib.or_ref_gapless_select(n_q, off.ctypes.data, *[a.ctypes.data for a in arrs], top_k, min_per_vid,
                                       adaptive, keep.ctypes.data, ks.ctypes.data, cal)
        dup = C.cast(C.addressof(cal) + 8, C.POINTER(C.c_float)).contents.value
        return keep[:n].tolist(), ks[:n_q].tolist(), (cal[0], cal[1], dup)

    def index(self, genome, ctg_off, ctg_len, min_aa_len=30, kmer=6, mod_bit=1, bbit=8):
        """the reference's k-mer block index (oracle/idx_ref.c): (ki, kb, bo) as numpy arrays"""
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        h = self.lib.or_ref_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, min_aa_len, kmer,
                                        mod_bit, bbit)
        try:
            n_kb = self.lib.or_ref_index_get(h, None, None, None)
            ki = np.zeros(1 << (4 * kmer - mod_bit), np.int64)
            kb = np.zeros(max(1, n_kb), np.uint32)
            bo = np.zeros(2 * len(off) + 1, np.uint32)
            self.lib.or_ref_index_get(h, ki.ctypes.data, kb.ctypes.data, bo.ctypes.data)
        finally:
            self.lib.or_ref_index_free(h)
        return ki, kb[:n_kb], bo

    def anchors(self, genome, ctg_off, ctg_len, queries, min_aa_len=30, kmer=6, mod_bit=1, bbit=8, max_occ=20000):
        """the reference's anchors of each query (ASCII bytes) against the genome's index
        (oracle/map_ref.c, oracle/idx_ref.c): one sorted numpy uint64 array per query"""
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        h = self.lib.or_ref_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, min_aa_len, kmer,
                                        mod_bit, bbit)
        out = []
        try:
            for q in queries:
                n = self.lib.or_ref_anchors(h, bytes(q), len(q), max_occ, 0, None)
                a = np.zeros(max(1, n), np.uint64)
                self.lib.or_ref_anchors(h, bytes(q), len(q), max_occ, n, a.ctypes.data)
                out.append(a[:n])
        finally:
            self.lib.or_ref_index_free(h)
        return out

    def chain(self, anchors, max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, coef, is_spliced, kmer,
              bbit):
        """the reference's mp_chain on one query's sorted anchors: (u list, anchors list)"""
        a = np.ascontiguousarray(anchors, dtype=np.uint64)
        n = len(a)
        u = np.zeros(max(1, n), np.uint64)
        out = np.zeros(max(1, n), np.uint64)
        n_u = self.lib.or_ref_chain(max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, coef, is_spliced,
                                    kmer, bbit, n, a.ctypes.data if n else None, len(u), u.ctypes.data, len(out),
                                    out.ctypes.data)
        m = int(sum(int(x) & 0xffffffff for x in u[:n_u]))
        return [int(x) for x in u[:n_u]], [int(x) for x in out[:m]]

    def orfs(self, ctg_nt4, rev, min_orf):
        """ps_orf_worker on strand rev of a contig: [(start, codons)], residues"""
        g = np.ascontiguousarray(ctg_nt4, dtype=np.uint8)
        cap = len(g) + 8
        st, ln, aa = np.zeros(cap, np.int64), np.zeros(cap, np.int32), np.zeros(cap, np.uint8)
        tot = C.c_int64()
        n = self.lib.or_ref_orfs(g.ctypes.data, len(g), rev, min_orf, cap, st.ctypes.data, ln.ctypes.data, cap,
                                 aa.ctypes.data, C.byref(tot))
        return [(int(st[i]), int(ln[i])) for i in range(n)], aa[:tot.value].tolist()

    def output(self, ctg_name, ctg_nt4, qname, aa, rev, qs, qe, vs, ve, as_, ae, cigar, score, o, pssm_row=None,
               profile_score=-2 ** 31, rank_sc=0, chn_sc=0, flag=0, acc_flags=0, gff_prefix=b"PL", id=0,
               gff_delim=-1, hit_idx=0):
        """the reference's mp_extra_gen/stop/start/cal + mp_write_output for one alignment (oracle/aln_ref.c)"""
        g = np.ascontiguousarray(ctg_nt4, dtype=np.uint8)
        cig = np.ascontiguousarray(cigar, dtype=np.uint32)
        mat = (C.c_int8 * 484).from_address(o.sc)
        prow = None if pssm_row is None else np.ascontiguousarray(pssm_row, dtype=np.int8)
        args = (ctg_name, g.ctypes.data, len(g), qname, aa, len(aa), rev, qs, qe, vs, ve, as_, ae,
                cig.ctypes.data, len(cig), score, profile_score, o.go, o.ge, o.fs, C.addressof(mat),
                None if prow is None else prow.ctypes.data, flag, acc_flags, rank_sc, chn_sc, gff_prefix, id,
                gff_delim, hit_idx)
        n = self.lib.or_ref_output(*args, None, 0)
        assert n >= 0
        buf = C.create_string_buffer(n + 1)
        self.lib.or_ref_output(*args, buf, n + 1)
        return buf.raw[:n]


def have_ref():
    return os.path.exists(REF_SO)


class ExtResult(C.Structure):
    _fields_ = [("nt_len", C.c_int32), ("aa_len", C.c_int32), ("score", C.c_int32)]


class Product(_Lib):
    """the product library: GPU kernels behind the reference's ns_* C-ABI"""

    def __init__(self, path=PRODUCT_SO):
        self.lib = L = C.CDLL(path)
        L.ns_make_tables.argtypes = [C.c_int]
        L.ns_opt_init.argtypes = [C.c_void_p]
        L.ns_opt_set_sp.argtypes = [C.c_void_p, C.c_int32]
        for f in (L.ns_global_gs16b, L.ns_global_gs32b, L.ns_global_gs_auto):
            f.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_void_p]
        L.ns_gpu_init.argtypes = [C.c_int32]
        L.ns_gpu_error.restype = C.c_char_p
        L.ns_gpu_name.restype = C.c_char_p
        L.ns_ext_stripe_width.argtypes = [C.c_void_p, C.c_int32]
        L.ns_ext_ie_penalty.argtypes = [C.c_void_p, C.c_int32]
        L.ns_ext_batch_init.argtypes = [C.c_void_p]
        L.ns_ext_batch_init.restype = C.c_void_p
        L.ns_ext_batch_add.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p]
        L.ns_ext_batch_flush.argtypes = [C.c_void_p]
        L.ns_ext_batch_ready.argtypes = [C.c_void_p, C.c_int32]
        L.ns_ext_batch_wait.argtypes = [C.c_void_p]
        L.ns_ext_batch_get.argtypes = [C.c_void_p, C.c_int32]
        L.ns_ext_batch_get.restype = C.POINTER(NsRst)
        L.ns_ext_batch_destroy.argtypes = [C.c_void_p]
        L.ns_ext_gpu_request_count.restype = C.c_uint64
        L.ns_ext_gpu_launch_count.restype = C.c_uint64
        L.ns_gpu_dispatch_count.restype = C.c_uint64
        L.ns_gpu_launch_count.restype = C.c_uint64
        L.nsh_ext_plan_create.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                          C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p]
        L.nsh_ext_plan_create.restype = C.c_void_p
        for f in (L.nsh_ext_plan_run, L.nsh_ext_plan_sync, L.nsh_ext_plan_destroy):
            f.argtypes = [C.c_void_p]
        L.nsh_ext_plan_run_subset.argtypes = [C.c_void_p, C.c_char_p]
        L.nsh_ext_plan_run_graph.argtypes = [C.c_void_p]
        L.nsh_ext_plan_pair_rows.argtypes = [C.c_void_p]
        L.nsh_ext_plan_results.argtypes = [C.c_void_p, C.c_void_p]
        L.nsh_ext_plan_cells.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_ext_plan_cells.restype = C.c_double
        L.nsh_ext_plan_last_ms.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_ext_plan_last_ms.restype = C.c_float
        L.nsh_ext_plan_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.nsh_ext_plan_autotune.restype = C.c_int
        L.nsh_build_info.restype = C.c_char_p
        L.ns_gpu_global.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p,
                                    C.c_void_p]
        L.ns_gpu_submit.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p]
        L.ns_gpu_submit.restype = C.c_void_p
        L.ns_gpu_wait.argtypes = [C.c_void_p, C.c_void_p]
        L.nsh_hmm_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        L.nsh_hmm_create.restype = C.c_void_p
        L.nsh_hmm_score_batch.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.nsh_hmm_destroy.argtypes = [C.c_void_p]
        L.nsh_pssm_ungapped_batch.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ns_gpu_gapless_score_batch.argtypes = [C.c_char_p, C.c_int64, C.c_char_p, C.c_int64, C.c_void_p, C.c_int32,
                                                 C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                 C.c_int32, C.c_char_p, C.c_void_p]
        L.ns_gpu_gapless_profile.argtypes = [C.c_void_p]
        assert L.ns_make_tables(1) == 0
        self._opt_init = L.ns_opt_init
        self._opt_set_sp = L.ns_opt_set_sp
        self.free = C.CDLL(None).free
        self.free.argtypes = [C.c_void_p]

    def error(self):
        return self.lib.ns_gpu_error().decode()

    def run(self, fn, ns, as_, opt, ss=None):
        r = NsRst()
        a = as_.encode() if isinstance(as_, str) else bytes(as_)
        b_ss, _ = _buf(ss)
        getattr(self.lib, fn)(None, bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, C.byref(r))
        out = (r.nt_len, r.aa_len, r.score, cigar_str(r))
        if r.cigar:
            self.free(C.cast(r.cigar, C.c_void_p))
        return out

    def gpu_global(self, ns, as_, opt, ss=None):
        """ns_gpu_global: (ok, (nt_len, aa_len, score, cigar))"""
        r = NsRst()
        a = as_.encode() if isinstance(as_, str) else bytes(as_)
        b_ss, _ = _buf(ss)
        ok = self.lib.ns_gpu_global(None, bytes(ns), len(ns), a, len(a), C.byref(opt), b_ss, C.byref(r))
        out = (r.nt_len, r.aa_len, r.score, cigar_str(r))
        if r.cigar:
            self.free(C.cast(r.cigar, C.c_void_p))
        return ok, out

    def ext_plan(self, ns, ns_off, nl, as_, as_off, al, opt, ss=None, dirs=None):
        nl = np.ascontiguousarray(nl, dtype=np.int32)
        al = np.ascontiguousarray(al, dtype=np.int32)
        ns_off = np.ascontiguousarray(ns_off, dtype=np.int64)
        as_off = np.ascontiguousarray(as_off, dtype=np.int64)
        d = None if dirs is None else bytes(np.ascontiguousarray(dirs, dtype=np.uint8))
        p = self.lib.nsh_ext_plan_create(len(nl), ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data,
                                         al.ctypes.data, ss, d, C.byref(opt))
        if not p:
            raise RuntimeError("nsh_ext_plan_create: " + self.error())
        return ExtPlan(self, p, len(nl))


class ExtPlan:
    def __init__(self, prod, p, n):
        self.prod, self.p, self.n = prod, p, n

    def run(self, sync=True, graph=False):
        L = self.prod.lib
        if not (L.nsh_ext_plan_run_graph(self.p) if graph else L.nsh_ext_plan_run(self.p)):
            raise RuntimeError("nsh_ext_plan_run%s: %s" % ("_graph" if graph else "", self.prod.error()))
        if sync and not L.nsh_ext_plan_sync(self.p):
            raise RuntimeError("nsh_ext_plan_sync: " + self.prod.error())

    def results(self):
        out = (ExtResult * self.n)()
        if not self.prod.lib.nsh_ext_plan_results(self.p, out):
            raise RuntimeError("nsh_ext_plan_results: " + self.prod.error())
        return [(r.nt_len, r.aa_len, r.score) for r in out]

    def cells(self, kind):
        return self.prod.lib.nsh_ext_plan_cells(self.p, kind)

    def ms(self, which):
        return self.prod.lib.nsh_ext_plan_last_ms(self.p, which)

    def pair_rows(self):
        """the row limit up to which small problems share a wave (2**31 - 1: all, 0: none)"""
        return self.prod.lib.nsh_ext_plan_pair_rows(self.p)

    def autotune(self, reps=2):
        """the persistent wave count per column-count group and the wavefronts per problem, picked
        by timing this plan's runs"""
        w = (C.c_int32 * 3)()
        if not self.prod.lib.nsh_ext_plan_autotune(self.p, reps, w):
            raise RuntimeError("nsh_ext_plan_autotune: " + self.prod.error())
        return list(w)

    def close(self):
        if self.p:
            self.prod.lib.nsh_ext_plan_destroy(self.p)
            self.p = None
