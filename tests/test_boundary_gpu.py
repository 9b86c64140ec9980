"""The drop-in boundary on the GPU:

  * the reference's GPU profile-HMM entry points (mp_hmm_dp_score_gpu_batch, mp_hmm_dp_score_gpu,
    mp_hmm_gpu_release; mppriv.h:128-134) reached the way hmm.c / map.c reach them: the profile is
    the caller's object, read through the caller's accessors (tests/hmm_caller.c stands in for
    hmm.c), its device copy cached in the caller's slot and freed by the release call. Scores are
    the golden ones the reference's mp_hmm_dp_score wrote;
  * the reference's own program (oracle/_ref/protlocus-hip: main.c, map.c, align.c, hmm.c, ...
    compiled from its sources and linked against libnasw_hip.so in place of its DP objects) run on
    the inputs of its integration tests under several option sets, with and without --gpu (the
    batched extension API and ns_gpu_submit/wait), and --gpu-min-cells=0 (every gap DP through
    ns_gpu_submit). Its output must equal, byte for byte, what the reference's CPU program wrote
    (golden/dropin_cases.json.gz; the first is the published DPP3 output, SHA-256 d303e5f3...).
"""
import base64
import ctypes as C
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import nsffi

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_hmm_entry_points_through_caller_accessors(oracle, product):
    so = os.path.join(HERE, "libhmm_caller.so")
    if not os.path.exists(so):
        raise RuntimeError("tests/libhmm_caller.so is not built (run __graft_entry__.build())")
    hc = C.CDLL(so, mode=C.RTLD_GLOBAL)
    hc.hc_make.restype = C.c_void_p
    hc.hc_make.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    hc.hc_slot0.restype = C.c_void_p
    hc.hc_slot0.argtypes = [C.c_void_p]
    hc.hc_slot_calls.argtypes = [C.c_void_p]
    hc.hc_free.argtypes = [C.c_void_p]
    L = product.lib
    L.mp_hmm_dp_score_gpu_batch.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
    L.mp_hmm_dp_score_gpu.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
    L.mp_hmm_dp_score_gpu.restype = C.c_int32
    L.mp_hmm_gpu_release.argtypes = [C.c_void_p]
    with gzip.open(os.path.join(HERE, "golden", "hmm_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for case in cases:
        h = nsffi.OrHmm()
        assert oracle.lib.or_hmm_parse(case["hmm"].encode(), C.byref(h), None) == 1
        M = h.len
        emit = np.ctypeslib.as_array(h.emit, shape=(M * 20,)).copy()
        trans = np.ctypeslib.as_array(h.trans, shape=(M * 7,)).copy()
        oracle.lib.or_hmm_free(C.byref(h))
        hh = hc.hc_make(M, emit.ctypes.data, trans.ctypes.data)
        regions = [bytes.fromhex(r) for r in case["regions"]]
        nt = b"".join(regions)
        off = np.cumsum([0] + [len(r) for r in regions[:-1]]).astype(np.int32)
        ln = np.array([len(r) for r in regions], dtype=np.int32)
        out = np.zeros(len(regions), dtype=np.int32)
        assert hc.hc_slot0(hh) is None
        assert L.mp_hmm_dp_score_gpu_batch(hh, len(regions), nt, max(1, len(nt)), off.ctypes.data, ln.ctypes.data,
                                           out.ctypes.data), product.error()
        assert out.tolist() == case["scores"], M
        assert hc.hc_slot0(hh) is not None                    # the device profile is cached in the caller's slot
        for r, exp in zip(regions, case["scores"]):
            got = L.mp_hmm_dp_score_gpu(hh, r, len(r))
            assert got == (exp if len(r) >= 3 else -2 ** 31), (M, len(r))
        L.mp_hmm_gpu_release(hh)
        assert hc.hc_slot0(hh) is None
        L.mp_hmm_gpu_release(hh)                             # releasing twice is harmless
        hc.hc_free(hh)


def test_reference_program_on_the_library():
    from golden.make_golden import DROPIN_INPUTS
    exe = os.path.join(ROOT, "oracle", "_ref", "protlocus-hip")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/protlocus-hip was not built (needs the reference tree at build time)")
    with gzip.open(os.path.join(HERE, "golden", "dropin_cases.json.gz"), "rt") as f:
        cases = json.load(f)
    for c in cases:
        g, p = (os.path.join(HERE, "golden", "dropin", f) for f in DROPIN_INPUTS[c["input"]])
        for extra in ([], ["--gpu"], ["--gpu", "--gpu-min-cells=0"]):
            args = extra + c["args"]
            print("drop-in run:", c["input"], " ".join(args), flush=True)
            r = subprocess.run([exe] + args + [g, p], capture_output=True, timeout=120)
            assert r.returncode == 0, (args, r.stderr.decode()[-1500:])
            assert b"GPU unavailable" not in r.stderr, r.stderr.decode()[-800:]
            assert r.stdout == base64.b64decode(c["stdout"]), (c["input"], args, r.stdout[:400])
