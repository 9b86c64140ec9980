"""nsh_ext_run_noalloc, the extension entry point that allocates no device memory (its buffers are
carved from one fixed device array per GPU), against the plan entry points it stands in for
(nsh_ext_plan_create + run + results) on the same inputs: bit-exact results on two calls in a
row, the second with a smaller input (the array keeps the first call's contents), with mixed
directions and the plain score-only mode; inputs beyond its stated limits are refused with a
message."""
import ctypes as C

import numpy as np
import pytest

import nsffi
from nsffi import NS_F_EXT_RIGHT

pytestmark = pytest.mark.gpu


def _noalloc(product, pairs, dirs, opt):
    from protlocus_amd import synth
    L = product.lib
    L.nsh_ext_run_noalloc.argtypes = [C.c_int32, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                      C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.nsh_ext_run_noalloc.restype = C.c_int
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    ns_off = np.ascontiguousarray(ns_off, dtype=np.int64)
    as_off = np.ascontiguousarray(as_off, dtype=np.int64)
    nl = np.ascontiguousarray(nl, dtype=np.int32)
    al = np.ascontiguousarray(al, dtype=np.int32)
    out = (nsffi.ExtResult * len(pairs))()
    d = None if dirs is None else bytes(np.ascontiguousarray(dirs, dtype=np.uint8))
    ok = L.nsh_ext_run_noalloc(len(pairs), ns, ns_off.ctypes.data, nl.ctypes.data, as_, as_off.ctypes.data,
                               al.ctypes.data, d, C.byref(opt), out, None)
    return ok, [(r.nt_len, r.aa_len, r.score) for r in out]


def _plan(product, pairs, dirs, opt):
    from protlocus_amd import synth
    ns, ns_off, nl, as_, as_off, al = synth.pack_batch(pairs)
    plan = product.ext_plan(ns, ns_off, nl, as_, as_off, al, opt, dirs=dirs)
    try:
        plan.run()
        return plan.results()
    finally:
        plan.close()


def test_noalloc_matches_plan_twice(product):
    from protlocus_amd import synth
    rng = np.random.default_rng(3)
    for n, lo, hi, mode in ((600, 50, 1000, "dirs"), (97, 20, 300, "right"), (40, 30, 200, "global")):
        print("no-alloc call:", n, lo, hi, mode, flush=True)
        pairs = synth.extension_batch(n, seed=int(rng.integers(1 << 30)), min_len=lo, max_len=hi)
        dirs = rng.integers(0, 2, size=n).astype(np.uint8) if mode == "dirs" else None
        opt = product.opt(flag=0 if mode == "global" else NS_F_EXT_RIGHT)
        ok, got = _noalloc(product, pairs, dirs, opt)
        assert ok, product.error()
        assert got == _plan(product, pairs, dirs, opt), (n, mode)


def test_noalloc_refuses_inputs_beyond_its_limits(product):
    from protlocus_amd import synth
    big = [(b"ACGT" * 40000, b"W" * 600)] * 240            # ~38 MB of DNA, records ~190 MB > 128 MB
    ok, _ = _noalloc(product, big, None, product.opt(flag=NS_F_EXT_RIGHT))
    assert not ok and "scratch" in product.error()
    long_prot = [(b"TGG" * 3000, b"W" * 2100)]                # padded protein 2112 > 2048
    ok, _ = _noalloc(product, long_prot, None, product.opt(flag=NS_F_EXT_RIGHT))
    assert not ok and "2048" in product.error()
    ok, got = _noalloc(product, [(b"TGG" * 19, b"W" * 19)], None, product.opt(flag=NS_F_EXT_RIGHT))
    assert ok and got == [(57, 19, 19 * 11 + 5)]          # still usable after a refusal
