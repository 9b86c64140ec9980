"""ctypes view of libnasw_hip.so for Python drivers (the prescreen trainer, the benchmark).

Types follow nasw.h:62-82; the engine entry points are declared in include/nasw-hip.h. Loading
fails loudly when the library is missing; every call that fails raises with ns_gpu_error().
"""
import ctypes as C
import os

import numpy as np

from . import LIB_PATH

NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT, NS_F_GPU = 1, 2, 4, 8
N_FEATURES = 64


class NsOpt(C.Structure):
    _fields_ = [("flag", C.c_int32), ("go", C.c_int32), ("ge", C.c_int32), ("io", C.c_int32), ("fs", C.c_int32),
                ("xdrop", C.c_int32), ("end_bonus", C.c_int32), ("asize", C.c_int32), ("gpu_device", C.c_int32),
                ("gpu_min_cells", C.c_int64), ("sp", C.c_int32 * 6), ("sp_null_bonus", C.c_int32),
                ("ie_coef", C.c_float), ("sc", C.c_void_p), ("pssm", C.c_void_p),
                ("nt4", C.c_void_p), ("aa20", C.c_void_p), ("codon", C.c_void_p)]


class NsRst(C.Structure):
    _fields_ = [("n_cigar", C.c_int32), ("m_cigar", C.c_int32), ("nt_len", C.c_int32), ("aa_len", C.c_int32),
                ("score", C.c_int32), ("cigar", C.POINTER(C.c_uint32))]


class Aln(C.Structure):
    """nsh_aln_t: one alignment as the output writers see it"""
    _fields_ = [("qname", C.c_char_p), ("aa", C.c_char_p), ("qlen", C.c_int32), ("qs", C.c_int32), ("qe", C.c_int32),
                ("ctg_name", C.c_char_p), ("ctg_len", C.c_int64), ("rev", C.c_int32), ("vs", C.c_int64),
                ("ve", C.c_int64), ("dp_score", C.c_int32), ("profile_score", C.c_int32), ("rank_sc", C.c_int32),
                ("chn_sc", C.c_int32), ("chn_sc_ungap", C.c_int32), ("cnt", C.c_int32),
                ("cigar", C.c_void_p), ("n_cigar", C.c_int32)]


class AlnStats(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("dp_max", "blen", "n_fs", "n_stop", "n_iden", "n_plus", "dist_stop",
                                          "dist_start", "n_exon", "n_feat")]


class Feat(C.Structure):
    _fields_ = [("vs", C.c_int64), ("ve", C.c_int64), ("qs", C.c_int32), ("qe", C.c_int32), ("type", C.c_int16),
                ("phase", C.c_int16), ("n_fs", C.c_int32), ("n_stop", C.c_int32), ("score", C.c_int32),
                ("n_iden", C.c_int32), ("blen", C.c_int32), ("donor", C.c_char * 2), ("acceptor", C.c_char * 2)]


MP_F_GFF, MP_F_NO_PAF, MP_F_GTF, MP_F_NO_CS, MP_F_ACCURACY, MP_ACC_RERANK = 0x8, 0x10, 0x20, 0x200, 0x400, 0x8
INT32_MIN = -2 ** 31


class ExtResult(C.Structure):
    _fields_ = [("nt_len", C.c_int32), ("aa_len", C.c_int32), ("score", C.c_int32)]


_lib = None


def lib():
    """the product library, loaded once (raises when it is not built)"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnasw_hip.so is not built: run python -m protlocus_amd.build")
        L = C.CDLL(LIB_PATH)
        L.ns_make_tables.argtypes = [C.c_int]
        L.ns_opt_init.argtypes = [C.c_void_p]
        L.ns_gpu_init.argtypes = [C.c_int32]
        L.ns_gpu_error.restype = C.c_char_p
        L.ns_gpu_name.restype = C.c_char_p
        L.nsh_ext_plan_create.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                          C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p]
        L.nsh_ext_plan_create.restype = C.c_void_p
        for f in (L.nsh_ext_plan_run, L.nsh_ext_plan_sync, L.nsh_ext_plan_destroy):
            f.argtypes = [C.c_void_p]
        L.nsh_ext_plan_run_subset.argtypes = [C.c_void_p, C.c_char_p]
        L.nsh_ext_plan_features.argtypes = [C.c_void_p, C.c_void_p]
        L.nsh_ext_plan_features.restype = C.c_void_p
        L.nsh_ext_plan_results.argtypes = [C.c_void_p, C.c_void_p]
        L.nsh_ext_plan_cells.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_ext_plan_cells.restype = C.c_double
        L.nsh_ext_plan_last_ms.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_ext_plan_last_ms.restype = C.c_float
        L.nsh_ext_plan_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.nsh_ext_plan_autotune.restype = C.c_int
        L.nsh_mlp_param_count.restype = C.c_int32
        L.nsh_mlp_create.argtypes = [C.c_void_p]
        L.nsh_mlp_create.restype = C.c_void_p
        L.nsh_mlp_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_void_p]
        L.nsh_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.nsh_mlp_infer.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
        L.nsh_mlp_set_norm.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.nsh_mlp_last_ms.argtypes = [C.c_void_p]
        L.nsh_mlp_last_ms.restype = C.c_float
        L.nsh_mlp_destroy.argtypes = [C.c_void_p]
        L.nsh_features.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_void_p,
                                   C.c_void_p, C.c_int64, C.c_void_p]
        L.nsh_hmm_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        L.nsh_hmm_create.restype = C.c_void_p
        L.nsh_hmm_score_batch.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.nsh_hmm_last_ms.argtypes = [C.c_void_p]
        L.nsh_hmm_last_ms.restype = C.c_float
        L.nsh_hmm_destroy.argtypes = [C.c_void_p]
        L.nsh_aln_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int32]
        L.nsh_aln_stats.restype = C.c_int32
        L.nsh_write_output.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_int32, C.c_int32, C.c_char_p, C.c_int64, C.c_int32, C.c_int32]
        L.nsh_write_output.restype = C.c_int64
        L.nsh_chain_pssm_ev_batch.argtypes = [C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                              C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L.nsh_orf_scan.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                   C.c_void_p, C.c_void_p]
        L.nsh_orf_scan.restype = C.c_int64
        L.nsh_ps_gapped_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.nsh_prescreen_build.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        L.nsh_prescreen_build.restype = C.c_void_p
        L.nsh_prescreen_size.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_prescreen_size.restype = C.c_int64
        L.nsh_prescreen_search.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_void_p,
                                           C.c_int32]
        L.nsh_prescreen_search.restype = C.c_int32
        L.nsh_prescreen_search_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                                 C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                                 C.c_void_p, C.c_int64, C.c_void_p]
        L.nsh_prescreen_search_batch.restype = C.c_int64
        L.nsh_prescreen_stats.argtypes = [C.c_void_p, C.c_void_p]
        L.nsh_index_build.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4
        L.nsh_index_build.restype = C.c_void_p
        L.nsh_index_build_nt4.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4
        L.nsh_index_build_nt4.restype = C.c_void_p
        L.nsh_index_size.argtypes = [C.c_void_p, C.c_int32]
        L.nsh_index_size.restype = C.c_int64
        L.nsh_index_copy.argtypes = [C.c_void_p] * 4
        L.nsh_index_destroy.argtypes = [C.c_void_p]
        L.nsh_index_anchors.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_void_p, C.c_int64]
        L.nsh_index_anchors.restype = C.c_int64
        L.nsh_chain_batch.argtypes = ([C.c_int32, C.c_void_p, C.c_void_p] + [C.c_int32] * 7 +
                                      [C.c_float, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p])
        L.nsh_chain_batch.restype = C.c_int64
        L.nsh_index_chain.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                      C.POINTER(ChainOpt), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.c_int64]
        L.nsh_index_chain.restype = C.c_int64
        L.nsh_gapless_select_batch.argtypes = [C.c_int32] + [C.c_void_p] * 6 + [C.c_int32] * 3 + [C.c_void_p] * 3
        L.nsh_gapless_prefilter_batch.argtypes = ([C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                                                   C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                   C.c_int32, C.c_void_p, C.c_int32] + [C.c_void_p] * 5 +
                                                  [C.c_int32] * 3 + [C.c_void_p] * 4)
        L.nsh_prescreen_destroy.argtypes = [C.c_void_p]
        L.nsh_map_opt_default.argtypes = [C.c_void_p]
        L.nsh_mapper_init.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4
        L.nsh_mapper_init.restype = C.c_void_p
        L.nsh_map_batch.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_void_p]
        L.nsh_mapper_destroy.argtypes = [C.c_void_p]
        L.free.argtypes = [C.c_void_p]
        L.nsh_device_synchronize.restype = C.c_int
        L.nsh_build_info.restype = C.c_char_p
        if L.ns_make_tables(1) != 0:
            raise RuntimeError("ns_make_tables failed")
        _lib = L
    return _lib


def error():
    return lib().ns_gpu_error().decode()


def init_gpu(device=0):
    if not lib().ns_gpu_init(device):
        raise RuntimeError("GPU not usable: " + error())


def opt(**kw):
    o = NsOpt()
    lib().ns_opt_init(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def pack(seqs):
    """concatenate byte strings: (buffer, int64 offsets, int32 lengths)"""
    ln = np.array([len(x) for x in seqs], dtype=np.int32)
    off = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        off[1:] = np.cumsum(ln[:-1], dtype=np.int64)
    return b"".join(seqs), off, ln


class Plan:
    """a device-resident batch of DP problems (nsh_ext_plan_*)"""

    def __init__(self, dnas, prots, o, dirs=None):
        L = lib()
        self.ns, self.ns_off, self.nl = pack(dnas)
        self.as_, self.as_off, self.al = pack(prots)
        self.n = len(dnas)
        self._opt = o
        d = None if dirs is None else bytes(np.ascontiguousarray(dirs, dtype=np.uint8))
        self.p = L.nsh_ext_plan_create(self.n, self.ns, self.ns_off.ctypes.data, self.nl.ctypes.data, self.as_,
                                       self.as_off.ctypes.data, self.al.ctypes.data, None, d, C.byref(o))
        if not self.p:
            raise RuntimeError("nsh_ext_plan_create: " + error())

    def run(self, keep=None, sync=True):
        L = lib()
        ok = L.nsh_ext_plan_run(self.p) if keep is None else \
            L.nsh_ext_plan_run_subset(self.p, bytes(np.ascontiguousarray(keep, dtype=np.uint8)))
        if not ok:
            raise RuntimeError("plan run: " + error())
        if sync and not L.nsh_ext_plan_sync(self.p):
            raise RuntimeError("plan sync: " + error())

    def sync(self):
        if not lib().nsh_ext_plan_sync(self.p):
            raise RuntimeError("plan sync: " + error())

    def scores(self):
        out = (ExtResult * self.n)()
        if not lib().nsh_ext_plan_results(self.p, out):
            raise RuntimeError("plan results: " + error())
        return np.array([r.score for r in out], dtype=np.int64)

    def cells(self, kind=1):
        return lib().nsh_ext_plan_cells(self.p, kind)

    def ms(self, which=1):
        return lib().nsh_ext_plan_last_ms(self.p, which)

    def autotune(self, reps=2):
        """the persistent wave count per column-count group, picked by timing this plan's runs"""
        w = (C.c_int32 * 3)()
        if not lib().nsh_ext_plan_autotune(self.p, reps, w):
            raise RuntimeError("nsh_ext_plan_autotune: " + error())
        return list(w)

    def features_device(self):
        ptr = lib().nsh_ext_plan_features(self.p, None)
        if not ptr:
            raise RuntimeError("plan features: " + error())
        return ptr

    def features_host(self):
        X = np.zeros((self.n, N_FEATURES), dtype=np.float32)
        if not lib().nsh_ext_plan_features(self.p, X.ctypes.data):
            raise RuntimeError("plan features: " + error())
        return X

    def close(self):
        if self.p:
            lib().nsh_ext_plan_destroy(self.p)
            self.p = None


def features(prots, dnas):
    """the 64 prescreen features of (protein, DNA window) pairs, host arrays in, n x 64 out"""
    as_, as_off, al = pack(prots)
    ns, ns_off, nl = pack(dnas)
    out = np.zeros((len(prots), N_FEATURES), dtype=np.float32)
    if not lib().nsh_features(len(prots), as_, as_off.ctypes.data, al.ctypes.data, len(as_), ns,
                              ns_off.ctypes.data, nl.ctypes.data, len(ns), out.ctypes.data):
        raise RuntimeError("nsh_features: " + error())
    return out


class MLP:
    """the prescreen network (nsh_mlp_*): FP32 parameters in, BF16 MFMA kernels on the device"""

    def __init__(self, params, mu=None, sd=None):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == lib().nsh_mlp_param_count()
        self.h = lib().nsh_mlp_create(params.ctypes.data)
        if not self.h:
            raise RuntimeError("nsh_mlp_create: " + error())
        if mu is not None:
            mu = np.ascontiguousarray(mu, dtype=np.float32)
            sd = np.ascontiguousarray(sd, dtype=np.float32)
            if not lib().nsh_mlp_set_norm(self.h, mu.ctypes.data, sd.ctypes.data):
                raise RuntimeError("nsh_mlp_set_norm: " + error())

    def train_step(self, X, t, lr):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        if not lib().nsh_mlp_train_step(self.h, X.ctypes.data, t.ctypes.data, len(t), lr, C.byref(loss)):
            raise RuntimeError("nsh_mlp_train_step: " + error())
        return loss.value

    def params(self, grads=False):
        n = lib().nsh_mlp_param_count()
        p, g = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
        if not lib().nsh_mlp_get_params(self.h, p.ctypes.data, g.ctypes.data):
            raise RuntimeError("nsh_mlp_get_params: " + error())
        return (p, g) if grads else p

    def infer(self, X=None, n=None, device_ptr=None):
        if device_ptr is None:
            X = np.ascontiguousarray(X, dtype=np.float32)
            n = len(X)
        out = np.zeros(n, dtype=np.float32)
        ok = lib().nsh_mlp_infer(self.h, X.ctypes.data if device_ptr is None else device_ptr, n,
                                 0 if device_ptr is None else 1, out.ctypes.data)
        if not ok:
            raise RuntimeError("nsh_mlp_infer: " + error())
        return out

    def last_ms(self):
        return lib().nsh_mlp_last_ms(self.h)

    def close(self):
        if self.h:
            lib().nsh_mlp_destroy(self.h)
            self.h = None


class HMM:
    """a profile resident on the device (nsh_hmm_*): int16 emissions M x 20, transitions M x 7"""

    def __init__(self, emit, trans):
        self.emit = np.ascontiguousarray(emit, dtype=np.int16)
        self.trans = np.ascontiguousarray(trans, dtype=np.int16)
        self.M = len(self.emit)
        self.h = lib().nsh_hmm_create(self.M, self.emit.ctypes.data, self.trans.ctypes.data)
        if not self.h:
            raise RuntimeError("nsh_hmm_create: " + error())

    def score(self, nt, off, ln):
        """Viterbi scores of regions nt[off[i]:off[i]+ln[i]] (base codes)"""
        off32 = np.ascontiguousarray(off, dtype=np.int32)
        ln = np.ascontiguousarray(ln, dtype=np.int32)
        out = np.zeros(len(ln), dtype=np.int32)
        if not lib().nsh_hmm_score_batch(self.h, len(ln), nt, len(nt), off32.ctypes.data, ln.ctypes.data,
                                         out.ctypes.data):
            raise RuntimeError("nsh_hmm_score_batch: " + error())
        return out

    def last_ms(self):
        return lib().nsh_hmm_last_ms(self.h)

    def close(self):
        if self.h:
            lib().nsh_hmm_destroy(self.h)
            self.h = None


def alignment_output(qname, aa, qs, qe, ctg_name, ctg_len, rev, vs, ve, cigar, dp_score, nt, as_, o,
                     pssm_row=None, profile_score=INT32_MIN, rank_sc=0, chn_sc=0, flag=0, acc_flags=0,
                     gff_prefix=b"PL", id=0, gff_delim=-1, hit_idx=0):
    """statistics, features and records of one alignment (nsh_aln_stats + nsh_write_output).
    nt: oriented target base codes (0..4) of [as_, ae); cigar: sequence of len << 4 | op.
    Returns (n_feat, AlnStats or None, [Feat], text bytes)."""
    L = lib()
    cig = np.ascontiguousarray(cigar, dtype=np.uint32)
    ntb = np.ascontiguousarray(nt, dtype=np.uint8)
    a = Aln(qname=qname, aa=aa, qlen=len(aa), qs=qs, qe=qe, ctg_name=ctg_name, ctg_len=ctg_len, rev=rev, vs=vs, ve=ve,
            dp_score=dp_score, profile_score=profile_score, rank_sc=rank_sc, chn_sc=chn_sc, chn_sc_ungap=0, cnt=0,
            cigar=cig.ctypes.data if len(cig) else None, n_cigar=len(cig))
    st = AlnStats()
    feats = (Feat * 64)()
    prow = None if pssm_row is None else np.ascontiguousarray(pssm_row, dtype=np.int8)
    nf = L.nsh_aln_stats(C.byref(a), ntb.ctypes.data, as_, as_ + len(ntb), C.byref(o),
                         None if prow is None else prow.ctypes.data, C.byref(st), feats, 64)
    if nf < 0:
        raise RuntimeError("nsh_aln_stats: " + error())
    stp = C.byref(st) if nf > 0 else None
    n = L.nsh_write_output(None, 0, C.byref(a), stp, feats, ntb.ctypes.data, as_, flag, acc_flags, gff_prefix, id,
                           gff_delim, hit_idx)
    if n < 0:
        raise RuntimeError("nsh_write_output: " + error())
    buf = C.create_string_buffer(n + 1)
    L.nsh_write_output(buf, n + 1, C.byref(a), stp, feats, ntb.ctypes.data, as_, flag, acc_flags, gff_prefix, id,
                       gff_delim, hit_idx)
    return nf, (st if nf > 0 else None), list(feats[:max(nf, 0)]), buf.raw[:n]


def chain_pssm_ev(chains, kmer, min_seg, xdrop):
    """PSSM diagonal evidence of refine chains (nsh_chain_pssm_ev_batch). chains: list of
    (region base codes 0..4, qlen x 22 int8 PSSM, anchors x << 32 | q). Returns (ev per chain,
    [per-anchor (score, aa_end) or None])."""
    L = lib()
    n = len(chains)
    nts = [np.ascontiguousarray(c[0], dtype=np.uint8) for c in chains]
    pss = [np.ascontiguousarray(c[1], dtype=np.int8).ravel() for c in chains]
    ans = [np.ascontiguousarray(c[2], dtype=np.uint64) for c in chains]
    cat = lambda xs, t: np.concatenate(xs).astype(t) if xs and sum(len(x) for x in xs) else np.zeros(1, t)
    offs = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])[:-1]]).astype(np.int64) if xs else np.zeros(0, np.int64)
    nt, ps, an = cat(nts, np.uint8), cat(pss, np.int8), cat(ans, np.uint64)
    nt_off, ps_off, a_off = offs(nts), offs(pss), offs(ans)
    l_nt = np.array([len(x) for x in nts], dtype=np.int64)
    qlen = np.array([len(x) // 22 for x in pss], dtype=np.int32)
    n_a = np.array([len(x) for x in ans], dtype=np.int32)
    n_a_all = int(n_a.sum())
    ev = np.zeros(max(n, 1), dtype=np.int32)
    sc, se = np.zeros(max(n_a_all, 1), dtype=np.int32), np.zeros(max(n_a_all, 1), dtype=np.int32)
    ok = L.nsh_chain_pssm_ev_batch(n, nt.ctypes.data, int(l_nt.sum()), nt_off.ctypes.data, l_nt.ctypes.data,
                                   ps.ctypes.data, int(sum(len(x) for x in pss)), ps_off.ctypes.data, qlen.ctypes.data,
                                   an.ctypes.data, n_a_all, a_off.ctypes.data, n_a.ctypes.data, kmer, min_seg, xdrop,
                                   ev.ctypes.data, sc.ctypes.data, se.ctypes.data)
    if not ok:
        raise RuntimeError("nsh_chain_pssm_ev_batch: " + error())
    per = [None if s == INT32_MIN else (int(s), int(e)) for s, e in zip(sc[:n_a_all], se[:n_a_all])]
    out, o = [], 0
    for k in range(n):
        out.append(per[o:o + n_a[k]])
        o += n_a[k]
    return ev[:n].tolist(), out


def orf_scan(seq, min_orf):
    """six-frame ORF walk of one oriented sequence (nsh_orf_scan): [(start, codons)], residues"""
    L = lib()
    s = np.ascontiguousarray(seq, dtype=np.uint8)
    cap = len(s) + 8          # at most len / 3 runs and residues per frame
    st, ln, aa = np.zeros(cap, np.int64), np.zeros(cap, np.int32), np.zeros(cap, np.uint8)
    tot = C.c_int64()
    n = L.nsh_orf_scan(s.ctypes.data, len(s), min_orf, cap, st.ctypes.data, ln.ctypes.data, len(aa), aa.ctypes.data,
                       C.byref(tot))
    if n < 0:
        raise RuntimeError("nsh_orf_scan: " + error())
    return [(int(st[i]), int(ln[i])) for i in range(n)], aa[:tot.value].tolist()


def ps_gapped_batch(pssm, orfs, diags, band, go, ge):
    """banded local affine scores of one profile (L x 22 int8) against ORFs (residue code arrays)
    on their diagonals (nsh_ps_gapped_batch)"""
    L = lib()
    ps = np.ascontiguousarray(pssm, dtype=np.int8)
    buf, off, ln = pack([bytes(np.asarray(o, dtype=np.uint8)) for o in orfs])
    dg = np.ascontiguousarray(diags, dtype=np.int32)
    sw = np.zeros(max(len(orfs), 1), dtype=np.int32)
    b = np.frombuffer(buf, dtype=np.uint8) if buf else np.zeros(1, np.uint8)
    if not L.nsh_ps_gapped_batch(ps.ctypes.data, ps.shape[0], b.ctypes.data, len(buf), len(orfs), off.ctypes.data,
                                 ln.ctypes.data, dg.ctypes.data, band, go, ge, sw.ctypes.data):
        raise RuntimeError("nsh_ps_gapped_batch: " + error())
    return sw[:len(orfs)].tolist()


def pack_nt4(codes):
    """one code per base to the reference's 4-bit genome (mp_ntdb_t.seq): base i in the low
    nibble of byte i/2 when i is even, the high nibble when odd"""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    if len(c) % 2:
        c = np.append(c, np.uint8(0))
    return (c[0::2] & 15) | ((c[1::2] & 15) << 4)


def index_build(genome, ctg_off, ctg_len, min_aa_len=30, kmer=6, mod_bit=1, bbit=8, packed=False):
    """the k-mer block index built on the device (nsh_index_*): (ki, kb, bo) as numpy arrays;
    packed: genome is the 4-bit form of pack_nt4 (nsh_index_build_nt4)"""
    L = lib()
    g = np.ascontiguousarray(genome, dtype=np.uint8)
    off = np.ascontiguousarray(ctg_off, dtype=np.int64)
    ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
    build = L.nsh_index_build_nt4 if packed else L.nsh_index_build
    h = build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, min_aa_len, kmer, mod_bit, bbit)
    if not h:
        raise RuntimeError("nsh_index_build: " + error())
    try:
        n_kb, n_bucket = L.nsh_index_size(h, 0), L.nsh_index_size(h, 1)
        ki, kb = np.zeros(n_bucket, np.int64), np.zeros(max(1, n_kb), np.uint32)
        bo = np.zeros(2 * len(off) + 1, np.uint32)
        if not L.nsh_index_copy(h, ki.ctypes.data, kb.ctypes.data, bo.ctypes.data):
            raise RuntimeError("nsh_index_copy: " + error())
    finally:
        L.nsh_index_destroy(h)
    return ki, kb[:n_kb], bo


class ChainOpt(C.Structure):
    """nsh_chain_opt_t; the defaults are the reference's (mp_mapopt_init, options.c) for a
    spliced search: max_intron 200000, max_gap 1000, bw 200000, max_chn_max_skip 25,
    max_chn_iter 1000000, min_chn_cnt 3, min_chn_sc 0, chn_coef_log 0.75, pre-chain on"""
    _fields_ = [("max_dist_x", C.c_int32), ("max_dist_y", C.c_int32), ("bw", C.c_int32), ("max_skip", C.c_int32),
                ("max_iter", C.c_int32), ("min_cnt", C.c_int32), ("min_sc", C.c_int32), ("chn_coef_log", C.c_float),
                ("is_spliced", C.c_int32), ("pre_chain", C.c_int32)]

    def __init__(self, max_dist_x=200000, max_dist_y=1000, bw=200000, max_skip=25, max_iter=1000000, min_cnt=3,
                 min_sc=0, chn_coef_log=0.75, is_spliced=1, pre_chain=1):
        super().__init__(max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, chn_coef_log, is_spliced,
                         pre_chain)


class Index:
    """a k-mer block index resident on the device (nsh_index_*)"""

    def __init__(self, genome, ctg_off, ctg_len, min_aa_len=30, kmer=6, mod_bit=1, bbit=8):
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        self.p = lib().nsh_index_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, min_aa_len, kmer,
                                       mod_bit, bbit)
        if not self.p:
            raise RuntimeError("nsh_index_build: " + error())

    def anchors(self, queries, max_occ=20000):
        """the sorted anchors (block << 32 | query position) of each query (ASCII bytes)"""
        seq = b"".join(bytes(q) for q in queries)
        ln = np.array([len(q) for q in queries], dtype=np.int32)
        off = np.concatenate([[0], np.cumsum(ln)[:-1]]).astype(np.int64) if len(ln) else np.zeros(0, np.int64)
        aoff = np.zeros(len(ln) + 1, np.int64)
        cap = 24 * len(seq) + 1024   # a first guess; a second call when the anchors do not fit
        while True:
            a = np.empty(cap, np.uint64)
            n = lib().nsh_index_anchors(self.p, len(ln), seq, off.ctypes.data, ln.ctypes.data, max_occ,
                                        aoff.ctypes.data, a.ctypes.data, cap)
            if n < 0:
                raise RuntimeError("nsh_index_anchors: " + error())
            if n <= cap:
                return [a[aoff[q]:aoff[q + 1]] for q in range(len(ln))]
            cap = n

    def chain(self, queries, max_occ=20000, opt=None):
        """mp_map_coarse's anchors -> pre-chain -> chain on the device (nsh_index_chain): per query
        the chains (score << 32 | count) and their anchors, as numpy uint64 arrays"""
        opt = opt or ChainOpt()
        seq = b"".join(bytes(q) for q in queries)
        ln = np.array([len(q) for q in queries], dtype=np.int32)
        off = np.concatenate([[0], np.cumsum(ln)[:-1]]).astype(np.int64) if len(ln) else np.zeros(0, np.int64)
        uoff, coff = np.zeros(len(ln) + 1, np.int64), np.zeros(len(ln) + 1, np.int64)
        cu, ca = len(seq) + 1024, 8 * len(seq) + 1024   # first guesses; a second call when they do not fit
        while True:
            u, a = np.empty(cu, np.uint64), np.empty(ca, np.uint64)
            n = lib().nsh_index_chain(self.p, len(ln), seq, off.ctypes.data, ln.ctypes.data, max_occ, C.byref(opt),
                                      uoff.ctypes.data, u.ctypes.data, cu, coff.ctypes.data, a.ctypes.data, ca)
            if n < 0:
                raise RuntimeError("nsh_index_chain: " + error())
            if uoff[-1] <= cu and coff[-1] <= ca:
                return [(u[uoff[q]:uoff[q + 1]], a[coff[q]:coff[q + 1]]) for q in range(len(ln))]
            cu, ca = max(cu, int(uoff[-1])), max(ca, int(coff[-1]))

    def close(self):
        if self.p:
            lib().nsh_index_destroy(self.p)
            self.p = None


def chain_batch(anchor_lists, max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, coef, is_spliced,
                kmer, bbit):
    """mp_chain of every query's sorted anchors on the device (nsh_chain_batch): per query the
    chains (score << 32 | count) and their anchors, as numpy uint64 arrays"""
    lens = np.array([len(a) for a in anchor_lists], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a = (np.concatenate([np.asarray(x, dtype=np.uint64) for x in anchor_lists]) if len(anchor_lists)
         else np.zeros(0, np.uint64))
    a = np.ascontiguousarray(a, dtype=np.uint64)
    n_q = len(anchor_lists)
    n_u = np.zeros(max(1, n_q), np.int32)
    u, ao = np.zeros(max(1, len(a)), np.uint64), np.zeros(max(1, len(a)), np.uint64)
    r = lib().nsh_chain_batch(n_q, off.ctypes.data, a.ctypes.data, max_dist_x, max_dist_y, bw, max_skip, max_iter,
                              min_cnt, min_sc, coef, is_spliced, kmer, bbit, n_u.ctypes.data, u.ctypes.data,
                              ao.ctypes.data)
    if r < 0:
        raise RuntimeError("nsh_chain_batch: " + error())
    out = []
    for q in range(n_q):
        uq = u[off[q]:off[q] + n_u[q]]
        m = int((uq & np.uint64(0xffffffff)).sum()) if len(uq) else 0
        out.append((uq, ao[off[q]:off[q] + m]))
    return out


class GaplessCalib(C.Structure):
    """nsh_gapless_calib_t (mp_gapless_calib_t's layout)"""
    _fields_ = [("noise", C.c_int32), ("n_query", C.c_int32), ("dup_frac", C.c_float)]


def _sel_arrays(cand_off, qlen, chn, vid, vs):
    return [np.ascontiguousarray(x, dtype=t) for x, t in ((cand_off, np.int32), (qlen, np.int32), (chn, np.int32),
                                                           (vid, np.uint32), (vs, np.int64))]


def gapless_select(cand_off, qlen, scores, chn, vid, vs, top_k, min_per_vid, adaptive, as_numpy=False):
    """the --gpu-prefilter selection of a mini-batch (nsh_gapless_select_batch): keep flags, K per
    query, (noise, n_query, dup_frac); lists, or numpy arrays with as_numpy"""
    off, ql, ch, vi, v = _sel_arrays(cand_off, qlen, chn, vid, vs)
    sc = np.ascontiguousarray(scores, dtype=np.int32)
    n_q, n = len(off) - 1, int(off[-1])
    keep, ks = np.zeros(max(1, n), np.uint8), np.zeros(max(1, n_q), np.int32)
    cal = GaplessCalib()
    if not lib().nsh_gapless_select_batch(n_q, off.ctypes.data, ql.ctypes.data, sc.ctypes.data, ch.ctypes.data,
                                          vi.ctypes.data, v.ctypes.data, top_k, min_per_vid, adaptive,
                                          keep.ctypes.data, ks.ctypes.data, C.byref(cal)):
        raise RuntimeError("nsh_gapless_select_batch: " + error())
    if as_numpy:
        return keep[:n], ks[:n_q], (cal.noise, cal.n_query, cal.dup_frac)
    return keep[:n].tolist(), ks[:n_q].tolist(), (cal.noise, cal.n_query, cal.dup_frac)


def gapless_prefilter(packed_nt, n_bases, queries, cands, anchors, block_width, kmer, xdrop, mat, codon,
                      cand_off, qlen, chn, vid, vs, top_k, min_per_vid, adaptive, as_numpy=False):
    """gapless scores and the selection in one call (nsh_gapless_prefilter_batch): scores, keep
    flags, K per query, calibration (lists, or numpy arrays with as_numpy). cands / anchors: ctypes
    arrays of the ns_gpu_gapless_* structs"""
    off, ql, ch, vi, v = _sel_arrays(cand_off, qlen, chn, vid, vs)
    m = np.ascontiguousarray(mat, dtype=np.int8)
    n_q, n = len(off) - 1, len(cands)
    scores, keep, ks = np.zeros(n, np.int32), np.zeros(n, np.uint8), np.zeros(n_q, np.int32)
    cal = GaplessCalib()
    q = bytes(queries)
    cod = bytes(codon)
    asize = int(round(len(m) ** 0.5))
    if not lib().nsh_gapless_prefilter_batch(packed_nt, n_bases, q, len(q), cands, n, anchors, len(anchors),
                                             block_width, kmer, xdrop, m.ctypes.data, asize, cod, n_q,
                                             off.ctypes.data, ql.ctypes.data, ch.ctypes.data, vi.ctypes.data,
                                             v.ctypes.data, top_k, min_per_vid, adaptive, scores.ctypes.data,
                                             keep.ctypes.data, ks.ctypes.data, C.byref(cal)):
        raise RuntimeError("nsh_gapless_prefilter_batch: " + error())
    if as_numpy:
        return scores, keep, ks, (cal.noise, cal.n_query, cal.dup_frac)
    return scores.tolist(), keep.tolist(), ks.tolist(), (cal.noise, cal.n_query, cal.dup_frac)


class PsCand(C.Structure):
    """nsh_ps_cand_t (the layout of the reference's mp_cand1_t)"""
    _fields_ = [("vid", C.c_uint32), ("vs", C.c_int64), ("ve", C.c_int64), ("ev", C.c_double), ("aqs", C.c_int32),
                ("aqe", C.c_int32), ("asc", C.c_int32), ("avs", C.c_int64)]

    def tuple(self):
        return (self.vid, self.vs, self.ve, self.ev, self.aqs, self.aqe, self.asc, self.avs)


class ProfileSearch:
    """the profile prescreen search over one genome (nsh_prescreen_*): contigs of base codes 0..4"""

    def __init__(self, genome, ctg_off, ctg_len, kmer=6, min_orf=30):
        L = lib()
        g = np.ascontiguousarray(genome, dtype=np.uint8)
        off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        ln = np.ascontiguousarray(ctg_len, dtype=np.int64)
        self.p = L.nsh_prescreen_build(len(off), g.ctypes.data, off.ctypes.data, ln.ctypes.data, kmer, min_orf)
        if not self.p:
            raise RuntimeError("nsh_prescreen_build: " + error())

    def size(self, what):
        return lib().nsh_prescreen_size(self.p, what)

    def search(self, pssm, thr, kmer_max=65536, min_ungapped=15, max_seqs=4000, band=32, go=11, ge=1, max_evalue=0.0):
        ps = np.ascontiguousarray(pssm, dtype=np.int8).reshape(-1, 22)
        cap = 4096
        while True:
            out = (PsCand * cap)()
            n = lib().nsh_prescreen_search(self.p, ps.ctypes.data, ps.shape[0], thr, kmer_max, min_ungapped, max_seqs,
                                           band, go, ge, max_evalue, out, cap)
            if n < 0:
                raise RuntimeError("nsh_prescreen_search: " + error())
            if n <= cap:
                return [out[i].tuple() for i in range(n)]
            cap = n

    def search_batch(self, pssms, thr, kmer_max=65536, min_ungapped=15, max_seqs=4000, band=32, go=11, ge=1,
                     max_evalue=0.0):
        """the candidates of every profile (a list of L x 22 int8 PSSMs) in one batched search"""
        mats = [np.ascontiguousarray(m, dtype=np.int8).reshape(-1, 22) for m in pssms]
        lens = np.array([m.shape[0] for m in mats], dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate(mats) if mats else np.zeros((0, 22), np.int8))
        n_res = np.zeros(max(1, len(mats)), dtype=np.int32)
        cap = 4096
        while True:
            out = (PsCand * cap)()
            n = lib().nsh_prescreen_search_batch(self.p, len(mats), flat.ctypes.data, lens.ctypes.data, thr, kmer_max,
                                                 min_ungapped, max_seqs, band, go, ge, max_evalue, out, cap,
                                                 n_res.ctypes.data)
            if n < 0:
                raise RuntimeError("nsh_prescreen_search_batch: " + error())
            if n <= cap:
                break
            cap = n
        res, k = [], 0
        for c in n_res[:len(mats)]:
            res.append([out[i].tuple() for i in range(k, k + int(c))])
            k += int(c)
        return res

    def stats(self):
        """counters of the last search: windows, k-mers, index hits, double-hit diagonals, ORFs >=
        min_ungapped, aligned, kept"""
        st = np.zeros(7, dtype=np.int64)
        if not lib().nsh_prescreen_stats(self.p, st.ctypes.data):
            raise RuntimeError("nsh_prescreen_stats: " + error())
        keys = ("windows", "kmers", "index_hits", "double_hit_diagonals", "orfs_scored", "aligned", "kept")
        return dict(zip(keys, (int(v) for v in st)))

    def close(self):
        if self.p:
            lib().nsh_prescreen_destroy(self.p)
            self.p = None


class AlignOpt(C.Structure):
    """nsh_align_opt_t: the DP options of the map path (mp_mapopt_t's go .. mat)"""
    _fields_ = [("go", C.c_int32), ("ge", C.c_int32), ("io", C.c_int32), ("io_end", C.c_int32), ("fs", C.c_int32),
                ("xdrop", C.c_int32), ("end_bonus", C.c_int32), ("max_ext", C.c_int32), ("max_intron", C.c_int32),
                ("kmer2", C.c_int32), ("end_retry_win", C.c_int32), ("ie_coef", C.c_float), ("sp_scale", C.c_float),
                ("sp_model", C.c_int32), ("sp_null_bonus", C.c_int32), ("mat", C.c_int8 * 484)]


class MapOpt(C.Structure):
    """nsh_map_opt_t: mp_mapopt_t's fields the map path reads (defaults from nsh_map_opt_default)"""
    _fields_ = [("dp", AlignOpt), ("max_occ", C.c_int32), ("max_gap", C.c_int32), ("bw", C.c_int32),
                ("max_chn_max_skip", C.c_int32), ("max_chn_iter", C.c_int32), ("min_chn_cnt", C.c_int32),
                ("min_chn_sc", C.c_int32), ("chn_coef_log", C.c_float), ("max_ava", C.c_int32), ("best_n", C.c_int32),
                ("out_n", C.c_int32), ("mask_len", C.c_int32), ("pri_ratio", C.c_float), ("mask_level", C.c_float),
                ("out_sim", C.c_float), ("out_cov", C.c_float), ("flag", C.c_uint32), ("gff_delim", C.c_int32),
                ("gff_prefix", C.c_char_p), ("n_threads", C.c_int32), ("prescreen", C.c_void_p), ("ps_tau", C.c_float),
                ("ps_audit", C.c_float), ("ps_seed", C.c_uint64), ("cand_feat", C.c_void_p), ("cand_label", C.c_void_p),
                ("cand_cap", C.c_int64)]

    @classmethod
    def default(cls, **kw):
        o = cls()
        lib().nsh_map_opt_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


MP_F_GFF, MP_F_NO_PAF, MP_F_GTF, MP_F_NO_CS = 0x8, 0x10, 0x20, 0x200
MAP_STAGES = ("coarse chains", "region selection", "refine sketches", "refine chains", "align stage", "records")


class Mapper:
    """the map path on one GPU (nsh_mapper_*): a genome and its device index, proteins mapped in
    batches to the records mp_write_output writes (PAF, or GFF3 with MP_F_GFF)"""

    def __init__(self, ctg_names, ctg_seqs, min_aa_len=30, kmer=6, mod_bit=1, bbit=8):
        lut = np.full(256, 4, np.uint8)
        for i, c in enumerate(b"ACGT"):
            lut[c] = lut[c + 32] = i
        g = np.ascontiguousarray(lut[np.frombuffer(b"".join(ctg_seqs), np.uint8)])
        ln = np.array([len(x) for x in ctg_seqs], np.int64)
        names = (C.c_char_p * len(ctg_names))(*[n.encode() if isinstance(n, str) else n for n in ctg_names])
        self.p = lib().nsh_mapper_init(len(ctg_names), names, g.ctypes.data, ln.ctypes.data, min_aa_len, kmer,
                                       mod_bit, bbit)
        if not self.p:
            raise RuntimeError("nsh_mapper_init: " + error())
        self.id = C.c_int64(0)

    def map(self, names, seqs, opt=None):
        """records of one batch of proteins (bytes) and the call's 24 counters (nsh_map_batch)"""
        opt = opt or MapOpt.default()
        qn = (C.c_char_p * len(seqs))(*[n.encode() if isinstance(n, str) else n for n in names])
        qs = (C.c_char_p * len(seqs))(*[bytes(x) for x in seqs])
        ql = np.array([len(x) for x in seqs], np.int32)
        buf, n, st = C.c_void_p(), C.c_int64(), np.zeros(24, np.int64)
        if not lib().nsh_map_batch(self.p, len(seqs), qn, qs, ql.ctypes.data, C.byref(opt), C.byref(self.id),
                                   C.byref(buf), C.byref(n), st.ctypes.data):
            raise RuntimeError("nsh_map_batch: " + error())
        try:
            return C.string_at(buf, n.value), st
        finally:
            lib().free(buf)

    def close(self):
        if self.p:
            lib().nsh_mapper_destroy(self.p)
            self.p = None
