"""GPU parity on the single-pair workload (300-aa protein vs ~2.4-3 kb of DNA with 2 GT..AG
introns, one 1-nt indel and 500-nt flanks; seeds 0-7) under the CLI option set (mp_map2ns_opt of
the options.c defaults: fs 23, GENERIC splice model, stop scores -23 except stop/stop).

ns_global_gs_auto(NS_F_CIGAR) on the GPU must give the golden score and CIGAR that the
reference's nasw-sse.c wrote (golden/config0_cases.json.gz), and the library's writers must turn
that CIGAR into the reference's PAF and GFF3 records byte for byte. Fresh seeds are compared with
the oracle, which tests/test_config0_cpu.py pins on the same workload. The retry path of
mp_align_step (io_end 19 for the end exons, align.c:674-678) is covered with the same pairs as left
and right extensions at io 19 and 29.
"""
import base64
import ctypes as C
import gzip
import json
import os

import pytest

from cases import cli_mat, cli_opt_kw, nt4_codes, parse_cigar
from nsffi import NS_F_CIGAR, NS_F_EXT_LEFT, NS_F_EXT_RIGHT
from protlocus_amd import capi, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "config0_cases.json.gz")


def _cases():
    with gzip.open(GOLDEN, "rt") as f:
        return json.load(f)


def test_config0_cigar_and_records(oracle, product):
    for c in _cases():
        print("config0 seed", c["seed"], flush=True)
        dna, prot = base64.b64decode(c["dna"]), base64.b64decode(c["prot"])
        o = product.opt(**cli_opt_kw(oracle.blosum62, flag=NS_F_CIGAR))
        nl, al, score, cg = product.run("ns_global_gs_auto", dna, prot, o)
        assert (nl, al, score) == (len(dna), len(prot), c["score"]), c["seed"]
        cigar = parse_cigar(cg)
        assert cigar == c["cigar"], c["seed"]
        oo = capi.opt()
        mat = (C.c_int8 * 484)(*cli_mat(oracle.blosum62))
        oo.sc = C.cast(mat, C.c_void_p).value
        oo.fs = 23
        nt = nt4_codes(dna)
        for flag, key in ((0, "paf"), (0x8, "gff")):
            text = capi.alignment_output(b"q1", prot, 0, len(prot), b"chr1", len(dna), 0, 0, len(dna), cigar, score,
                                         nt, 0, oo, flag=flag)[3]
            assert text == base64.b64decode(c[key]), (c["seed"], key)


def test_config0_fresh_seeds_against_oracle(oracle, product):
    for seed in range(100, 108):
        dna, prot = synth.single_pair(seed)
        kw = cli_opt_kw(oracle.blosum62, flag=NS_F_CIGAR)
        exp = oracle.global_dp(dna, prot, oracle.opt(**kw), mode="auto")
        got = product.run("ns_global_gs_auto", dna, prot, product.opt(**kw))
        assert got == exp, seed


def test_config0_end_extensions_with_retry_io(oracle, product):
    """the end-exon extensions of mp_align_step with the CLI options, at io 29 and at io_end 19"""
    for seed in range(8):
        dna, prot = synth.single_pair(seed)
        for io in (29, 19):
            for flag in (NS_F_EXT_LEFT, NS_F_EXT_RIGHT):
                kw = cli_opt_kw(oracle.blosum62, flag=flag)
                kw["io"] = io
                exp = oracle.ext(dna, prot, oracle.opt(**kw))[0]
                got = product.run("ns_global_gs16b", dna, prot, product.opt(**kw))
                assert got[:3] == exp, (seed, io, flag)
