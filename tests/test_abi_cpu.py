"""CPU checks of the product library (no GPU): it loads, exports every function and table that
include/nasw-hip.h declares, and its host-side logic (tables, options, input preparation,
stripe width, length penalty) equals the oracle and, where built, the compiled reference.
"""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import nsffi
from nsffi import NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nasw-hip.h")

pytestmark = pytest.mark.skipif(not os.path.exists(nsffi.PRODUCT_SO),
                                reason="libnasw_hip.so not built (run __graft_entry__.build())")


def declared_symbols():
    """function and data names declared at file scope in the header (static inline excluded)"""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"static inline[^{]*\{(?:[^{}]|\{[^{}]*\})*\}", "", text)
    text = re.sub(r"typedef[^;]*;", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    funcs = set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text))
    data = set()
    for decl in re.findall(r"extern\s+(?:const\s+)?\w+\s+([^;(\"]+);", text):
        for item in decl.split(","):
            data.add(re.sub(r"\[[^\]]*\]", "", item).strip(" *"))
    return funcs - {"if", "while", "for", "return", "sizeof"}, data


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", nsffi.PRODUCT_SO], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_header_symbols_are_exported():
    funcs, data = declared_symbols()
    assert len(funcs) > 40 and "ns_global_gs16b" in funcs and "ns_tab_nt4" in data
    have = exported()
    missing = sorted((funcs | data) - have - {"kmalloc", "krealloc", "kcalloc", "kfree"})
    assert not missing, missing


def test_library_loads_and_tables_match():
    p = nsffi.Product()
    t = json.load(open(os.path.join(ROOT, "tests", "golden", "tables.json")))
    L = p.lib
    assert list((C.c_uint8 * 256).in_dll(L, "ns_tab_nt4")) == t["nt4"]
    assert list((C.c_uint8 * 256).in_dll(L, "ns_tab_aa20")) == t["aa20"]
    assert list((C.c_uint8 * 64).in_dll(L, "ns_tab_codon")) == t["codon"]
    assert list((C.c_int8 * 484).in_dll(L, "ns_mat_blosum62")) == t["blosum62"]
    assert L.nsh_build_info().decode().startswith("libnasw_hip")


def test_opt_init_matches_oracle(oracle):
    p = nsffi.Product()
    a, b = p.opt(), oracle.opt()
    for f in nsffi.OPT_FIELDS:
        assert getattr(a, f) == getattr(b, f), f
    assert list(a.sp) == list(b.sp)
    sa = (C.c_int8 * 484).from_address(a.sc)
    sb = (C.c_int8 * 484).from_address(b.sc)
    assert list(sa) == list(sb)


def test_ext_prep_matches_oracle(oracle):
    from cases import planted_case, random_ss
    p = nsffi.Product()
    L = p.lib
    L.ns_ext_prep.restype = C.c_void_p
    L.ns_ext_prep.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p,
                              C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(21)
    for k in range(60):
        ns, a = planted_case(rng, int(rng.integers(1, 120)))
        ss = random_ss(rng, len(ns), p=0.5)
        flag = NS_F_EXT_LEFT if k & 1 else NS_F_EXT_RIGHT
        nl, al = len(ns), len(a)
        aas, don, acc = C.c_void_p(), C.c_void_p(), C.c_void_p()
        blk = L.ns_ext_prep(None, ns, nl, a.encode(), al, C.byref(p.opt(flag=flag)), ss,
                            C.byref(aas), C.byref(don), C.byref(acc))
        assert blk
        got = (C.string_at(blk, nl), C.string_at(aas, al), C.string_at(don, nl + 1), C.string_at(acc, nl + 1))
        p.free(blk)
        nas_o = C.create_string_buffer(nl + 1)
        aas_o = C.create_string_buffer(al + 1)
        don_o = C.create_string_buffer(nl + 2)
        acc_o = C.create_string_buffer(nl + 2)
        oracle.lib.or_prep(ns, nl, a.encode(), al, C.byref(oracle.opt(flag=flag)), ss, 1 if k & 1 else 0,
                           nas_o, aas_o, don_o, acc_o)
        exp = (nas_o.raw[:nl], aas_o.raw[:al], don_o.raw[:nl + 1], acc_o.raw[:nl + 1])
        assert got == exp, k


def test_stripe_width_and_penalty(oracle):
    p = nsffi.Product()
    for flag in (0, NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT):
        for al in (1, 8, 15, 16, 17, 300):
            w = p.lib.ns_ext_stripe_width(C.byref(p.opt(flag=flag)), al)
            assert w in (8, 16)
            if nsffi.have_ref():
                r = nsffi.Reference()
                assert w == r.stripe_width(r.opt(flag=flag), al)
    o = p.opt()
    for d in list(range(0, 40)) + [100, 1000, 12345, 10 ** 6]:
        assert p.lib.ns_ext_ie_penalty(C.byref(o), d) == oracle.ie_penalty(o.ie_coef, d), d
