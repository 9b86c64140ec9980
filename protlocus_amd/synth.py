"""Seeded synthetic workloads for the spliced-DP benchmarks and parity tests.

The generators follow the workload descriptions the benchmarks are quoted on:
  * one pair: a protein drawn from BLOSUM62 background frequencies, reverse-translated with
    uniform synonymous codons, substituted along BLOSUM62 conditionals, split by GT..AG
    introns, with a 1-nt frameshift and random flanks;
  * extension batches: many such pairs with log-uniform protein lengths.
Background frequencies are the published BLOSUM62 ones (Henikoff & Henikoff 1992); the
conditionals P(b|a) are proportional to p_b * 2^(s_ab/2), i.e. the half-bit matrix read back
as target frequencies, excluding b == a.
"""
import math

import numpy as np

AA = "ARNDCQEGHILKMFPSTWYV"   # aa20 order of the C-ABI tables (first 20 codes)
BG = np.array([.074, .052, .045, .054, .025, .034, .054, .074, .026, .068,
               .099, .058, .025, .047, .039, .057, .051, .013, .032, .073])
BG = BG / BG.sum()
# standard genetic code in AAA, AAC, AAG, AAT, ACA, ... order
CODE1 = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
NT = np.frombuffer(b"ACGT", dtype=np.uint8)
_B62 = [
    [4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0],
    [-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3],
    [-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3],
    [-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3],
    [0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1],
    [-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2],
    [-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2],
    [0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3],
    [-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3],
    [-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3],
    [-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1],
    [-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2],
    [-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1],
    [-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1],
    [-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2],
    [1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2],
    [0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0],
    [-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3],
    [-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1],
    [0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4]]
B62 = np.array(_B62, dtype=np.float64)
_COND = BG[None, :] * np.power(2.0, B62 / 2.0)
np.fill_diagonal(_COND, 0.0)
COND = _COND / _COND.sum(axis=1, keepdims=True)
COND_CDF = np.cumsum(COND, axis=1)
SYN = {a: [i for i, c in enumerate(CODE1) if c == a] for a in AA}
CODON_STR = ["".join("ACGT"[(i >> s) & 3] for s in (4, 2, 0)) for i in range(64)]


def rand_dna(rng, n, gc=0.5):
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return NT[rng.choice(4, size=n, p=p)].tobytes()


_AAIDX = np.full(256, 0, dtype=np.int64)
for _i, _c in enumerate(AA):
    _AAIDX[ord(_c)] = _i
_SYN_TAB = np.zeros((20, 6), dtype=np.int64)
_SYN_N = np.zeros(20, dtype=np.int64)
for _i, _c in enumerate(AA):
    _SYN_N[_i] = len(SYN[_c])
    _SYN_TAB[_i, :len(SYN[_c])] = SYN[_c]
_CODON_NT = np.array([[ord(ch) for ch in c] for c in CODON_STR], dtype=np.uint8)
_AA_BYTES = np.frombuffer(AA.encode(), dtype=np.uint8)


def substitute(rng, prot, rate):
    codes = _AAIDX[np.frombuffer(prot, dtype=np.uint8)]
    hit = np.nonzero(rng.random(len(codes)) < rate)[0]
    if len(hit):
        u = rng.random(len(hit))
        cdf = COND_CDF[codes[hit]]
        codes = codes.copy()
        codes[hit] = np.minimum((cdf < (u * cdf[:, -1])[:, None]).sum(axis=1), 19)
    return _AA_BYTES[codes].tobytes()


def rand_protein(rng, n):
    return _AA_BYTES[rng.choice(20, size=n, p=BG)].tobytes()


def reverse_translate(rng, prot):
    codes = _AAIDX[np.frombuffer(prot, dtype=np.uint8)]
    pick = (rng.random(len(codes)) * _SYN_N[codes]).astype(np.int64)
    return _CODON_NT[_SYN_TAB[codes, pick]].tobytes()


def gene_model(rng, prot_len, sub_rate, n_introns, intron_len_fn, n_fs, flank):
    """protein + DNA carrying its (mutated) coding sequence; returns (dna, protein)"""
    prot = rand_protein(rng, prot_len)
    return gene_of(rng, prot, sub_rate, n_introns, intron_len_fn, n_fs, flank), prot


def gene_of(rng, prot, sub_rate, n_introns, intron_len_fn, n_fs, flank):
    """DNA carrying a mutated, spliced, frameshifted coding sequence of protein `prot`"""
    cds = reverse_translate(rng, substitute(rng, prot, sub_rate))
    seq = bytearray(cds)
    cuts = sorted(rng.integers(1, max(2, len(cds) - 1), size=n_introns).tolist(), reverse=True)
    for c in cuts:
        L = int(intron_len_fn())
        seq[c:c] = b"GT" + rand_dna(rng, max(0, L - 4)) + b"AG"
    for _ in range(n_fs):
        p = int(rng.integers(0, len(seq)))
        if rng.random() < 0.5:
            del seq[p]
        else:
            seq[p:p] = rand_dna(rng, 1)
    return rand_dna(rng, flank) + bytes(seq) + rand_dna(rng, flank)


def single_pair(seed=0):
    """one 300-aa protein vs ~2.4-3 kb DNA: 15% substitutions, 2 introns 80-400 nt, one 1-nt indel,
    500-nt flanks (configuration of the single-pair global DP check)"""
    rng = np.random.default_rng(seed)
    return gene_model(rng, 300, 0.15, 2, lambda: rng.integers(80, 401), 1, 500)


def extension_batch(n, seed=1, min_len=50, max_len=1000):
    """n (dna, protein) pairs: lengths log-uniform in [min_len, max_len], 10-30% substitutions,
    0-3 GT..AG introns (log-normal, median 300 nt, capped at 5 kb), 0-2 frameshifts, 200-nt flanks"""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(round(np.exp(rng.uniform(np.log(min_len), np.log(max_len)))))
        rate = rng.uniform(0.10, 0.30)
        ni = int(rng.integers(0, 4))
        nfs = int(rng.integers(0, 3))
        out.append(gene_model(rng, L, rate, ni, lambda: min(5000, max(8, rng.lognormal(np.log(300), 0.8))), nfs, 200))
    return out


def pack_batch(pairs):
    """concatenate pairs for nsh_ext_plan_create: (ns, ns_off, nl, as, as_off, al)"""
    nl = np.array([len(d) for d, _ in pairs], dtype=np.int32)
    al = np.array([len(p) for _, p in pairs], dtype=np.int32)
    ns_off = np.zeros(len(pairs), dtype=np.int64)
    as_off = np.zeros(len(pairs), dtype=np.int64)
    if len(pairs) > 1:
        ns_off[1:] = np.cumsum(nl.astype(np.int64))[:-1]
        as_off[1:] = np.cumsum(al.astype(np.int64))[:-1]
    return b"".join(d for d, _ in pairs), ns_off, nl, b"".join(p for _, p in pairs), as_off, al


# ---------------------------------------------------------------------------------------
# profile HMMs (HMMER3 text) and genomes for the profile workloads
# ---------------------------------------------------------------------------------------
HMM_ALPHA = "ACDEFGHIKLMNPQRSTVWY"   # HMMER3 emission column order
HMM_BG = np.array([BG[AA.index(c)] for c in HMM_ALPHA])


def hmm_params(rng, M, alpha=0.5, ins_ext=0.5, del_rate=0.05):
    """match emissions ~ Dirichlet(alpha * background) per state; i->i = ins_ext (expected
    insert length 1/(1-ins_ext)); m->d = del_rate. Returns probabilities (M x 20, M x 7) with
    transitions in HMMER order m->m m->i m->d i->m i->i d->m d->d."""
    em = rng.dirichlet(alpha * HMM_BG, size=M)
    em = np.maximum(em, 1e-6)
    em /= em.sum(axis=1, keepdims=True)
    mi = 0.02
    tr = np.zeros((M, 7))
    tr[:, 0], tr[:, 1], tr[:, 2] = 1 - mi - del_rate, mi, del_rate
    tr[:, 3], tr[:, 4] = 1 - ins_ext, ins_ext
    tr[:, 5], tr[:, 6] = 0.7, 0.3
    tr[-1, 2], tr[-1, 0] = 0.0, 1 - mi      # no delete out of the last state
    tr[-1, 5], tr[-1, 6] = 1.0, 0.0
    return em, tr


def hmm_text(em, tr, name="synthetic"):
    """one HMMER3 record: probabilities written as -ln(p) with 5 decimals, 0 as '*'"""
    M = len(em)
    f = lambda p: "*" if p <= 0 else "%.5f" % (-np.log(p))
    out = ["HMMER3/f [synthetic]", "NAME  %s" % name, "LENG  %d" % M, "ALPH  amino",
           "HMM          " + "        ".join(HMM_ALPHA),
           "            m->m     m->i     m->d     i->m     i->i     d->m     d->d",
           "  COMPO   " + " ".join(f(p) for p in HMM_BG),
           "          " + " ".join(f(p) for p in HMM_BG),
           "          " + " ".join(f(p) for p in [0.98, 0.01, 0.01, 0.5, 0.5, 1.0, 0.0])]
    for k in range(M):
        out.append("%7d   " % (k + 1) + " ".join(f(p) for p in em[k]) + " %d - - - -" % (k + 1))
        out.append("          " + " ".join(f(p) for p in HMM_BG))
        out.append("          " + " ".join(f(p) for p in tr[k]))
    out.append("//")
    return "\n".join(out) + "\n"


def hmm_sample_protein(rng, em, tr, max_len=None):
    """a protein emitted by a path through the profile (match, insert and delete states)"""
    M = len(em)
    seq, k, state = [], 0, 0   # state 0: match, 1: insert, 2: delete
    while k < M:
        if state == 0:
            seq.append(HMM_ALPHA[rng.choice(20, p=em[k])])
            r = rng.random()
            state = 0 if r < tr[k, 0] else (1 if r < tr[k, 0] + tr[k, 1] else 2)
            if state != 1:
                k += 1
        elif state == 1:
            seq.append(HMM_ALPHA[rng.choice(20, p=HMM_BG)])
            state = 1 if rng.random() < tr[k - 1 if k else 0, 4] else 0
            if state == 0:
                k += 1
        else:
            state = 2 if rng.random() < tr[k - 1 if k else 0, 6] else 0
            k += 1
    return "".join(seq[:max_len] if max_len else seq)


def hmm_tables(em, tr):
    """the int16 emission (M x 20, best residue at +4) and transition (M x 7) tables that the
    HMMER3 text of hmm_text() reads back as (-ln p printed with 5 decimals, score lrint(-2x))"""
    def score(p):
        p = np.asarray(p, dtype=np.float64)
        x = np.where(p > 0, np.round(-np.log(np.where(p > 0, p, 1.0)), 5), 0.0)
        v = np.clip(np.rint(-2.0 * x), -32767, 32767)
        return np.where(p > 0, v, -32768).astype(np.int64)
    e = score(em)
    e = e - e.max(axis=1, keepdims=True) + 4
    return e.astype(np.int16), score(tr).astype(np.int16)


# the log-odds background of the profile PSSM (HMMER's amino-acid frequencies, in HMMER column
# order), as the reference's hmm.c:28-31 holds it
HMMER_BG = [0.0787945, 0.0151600, 0.0535222, 0.0668298, 0.0397062, 0.0695071, 0.0229198, 0.0590092, 0.0594422,
            0.0963728, 0.0237718, 0.0414386, 0.0482904, 0.0395639, 0.0540978, 0.0683364, 0.0540687, 0.0673417,
            0.0114135, 0.0304133]


def hmm_pssm(em, stop_pen=23):
    """the L x 22 int8 PSSM that mp_hmm_read1 derives from hmm_text()'s record (hmm.c:69-81,
    152-161): lrint of 2 (-x - ln bg) / ln 2 from each printed -ln p = x, clamped to +-64 (-64 for
    '*'), in ns_tab_aa20 columns; stop column -stop_pen, X column -1"""
    aa20 = "ARNDCQEGHILKMFPSTWYV*X"
    P = np.zeros((len(em), 22), dtype=np.int8)
    for k, row in enumerate(em):
        for i, c in enumerate(HMM_ALPHA):
            p = row[i]
            if p <= 0:
                v = -64
            else:
                x = float("%.5f" % (-np.log(p)))
                lo = 2.0 * (-x - math.log(HMMER_BG[i])) / math.log(2.0)
                v = int(np.rint(min(max(lo, -64.0), 64.0)))
            P[k, aa20.index(c)] = v
        P[k, 20], P[k, 21] = -stop_pen, -1
    return P


def profile_genome(n_prof=64, genome_len=10_000_000, n_plant=200, seed=5, min_m=100, max_m=800):
    """n_prof profiles (M log-uniform) and a random genome (GC 0.45, base codes 0..3) carrying
    n_plant profile-sampled hits, reverse-translated, on either strand"""
    rng = np.random.default_rng(seed)
    profs = []
    for _ in range(n_prof):
        M = int(round(np.exp(rng.uniform(np.log(min_m), np.log(max_m)))))
        profs.append(hmm_params(rng, M))
    p = np.array([0.275, 0.225, 0.225, 0.275])
    g = rng.choice(4, size=genome_len, p=p).astype(np.uint8)
    lut = np.full(256, 4, dtype=np.uint8)
    lut[list(b"ACGT")] = [0, 1, 2, 3]
    for _ in range(n_plant):
        em, tr = profs[int(rng.integers(0, n_prof))]
        cds = lut[np.frombuffer(reverse_translate(rng, hmm_sample_protein(rng, em, tr).encode()), dtype=np.uint8)]
        if rng.random() < 0.5:
            cds = (3 - cds)[::-1]
        pos = int(rng.integers(0, genome_len - len(cds)))
        g[pos:pos + len(cds)] = cds
    return profs, g


def profile_scan(n_prof=64, genome_len=10_000_000, win=3000, step=2700, n_plant=200, seed=5,
                 min_m=100, max_m=800):
    """profile-HMM scan workload: n_prof profiles (M log-uniform), a random genome (GC 0.45) with
    n_plant profile-sampled hits (reverse-translated, either strand), cut into windows; each window
    is scored in its six frames (forward offsets 0-2, reverse complement offsets 0-2), one region
    per frame. Returns (profiles [(em, tr)], regions (bytes of base codes 0..3), offsets, lengths)."""
    profs, g = profile_genome(n_prof, genome_len, n_plant, seed, min_m, max_m)
    starts = np.arange(0, max(1, genome_len - win + 1), step)
    rc = (3 - g)[::-1]
    regs = []
    for s in starts:
        e = min(genome_len, s + win)
        fw = g[s:e]
        rv = rc[genome_len - e:genome_len - s]
        for f in range(3):
            regs.append(fw[f:])
            regs.append(rv[f:])
    lens = np.array([len(r) for r in regs], dtype=np.int32)
    offs = np.zeros(len(regs), dtype=np.int64)
    offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    return profs, np.concatenate(regs).tobytes(), offs, lens


_CODE_AA = np.array([AA.index(c) if c != "*" else 20 for c in CODE1], dtype=np.uint8)


def gapless_minibatch(n_query=5000, genome_len=20_000_000, n_contigs=24, seed=9, bw=256, n_anchor=8, kmer=6,
                      min_len=50, max_len=1000, min_cand=2, max_cand=48, sub_rate=0.2):
    """a --gpu-prefilter mini-batch (about 2 M query residues, the reference's default mini-batch):
    a random genome (GC 0.45) cut into contigs; per query (length log-uniform) its true locus, a
    frame of either strand translated with sub_rate substitutions, among min_cand..max_cand coarse
    candidates, the rest decoys at random; n_anchor anchors per candidate, spread along the true
    diagonal or at random for decoys, block-aligned as map.c:474-500 picks them. Returns a dict of
    numpy arrays: genome codes, queries (aa codes, concatenated), the candidate and anchor fields,
    and per candidate the chain score, vid and vs; per query cand_off and qlen."""
    rng = np.random.default_rng(seed)
    g = rng.choice(4, size=genome_len, p=[0.275, 0.225, 0.225, 0.275]).astype(np.uint8)
    cuts = np.sort(rng.choice(np.arange(100_000, genome_len - 100_000), n_contigs - 1, replace=False))
    ctg_off = np.concatenate([[0], cuts]).astype(np.int64)
    ctg_len = np.diff(np.concatenate([ctg_off, [genome_len]])).astype(np.int64)
    orient = {}

    def oriented(c, rev):
        if (c, rev) not in orient:
            o = g[ctg_off[c]:ctg_off[c] + ctg_len[c]]
            orient[(c, rev)] = (3 - o[::-1]) if rev else o
        return orient[(c, rev)]
    qlen = np.exp(rng.uniform(np.log(min_len), np.log(max_len), n_query)).astype(np.int32)
    p_ctg = ctg_len / ctg_len.sum()
    queries, q_off = [], [0]
    cand = {x: [] for x in ("target_offset", "target_length", "query_offset", "query_length", "anchor_offset",
                            "anchor_count", "reverse", "chn", "vid", "vs")}
    anch_block, anch_qend, cand_off = [], [], [0]
    for q in range(n_query):
        L = int(qlen[q])
        c, rev = int(rng.choice(n_contigs, p=p_ctg)), int(rng.integers(0, 2))
        o = oriented(c, rev)
        e0 = int(rng.integers(2, len(o) - 3 * L))
        cod = o[e0 - 2:e0 - 2 + 3 * L].reshape(L, 3)
        prot = _CODE_AA[cod[:, 0].astype(np.int64) << 4 | cod[:, 1] << 2 | cod[:, 2]]
        mut = rng.random(L) < sub_rate
        prot[mut] = rng.integers(0, 20, size=int(mut.sum()))
        queries.append(prot)
        n_c = int(rng.integers(min_cand, max_cand + 1))
        true_at = int(rng.integers(0, n_c))
        for j in range(n_c):
            if j == true_at:
                cc, rr, s0, chn = c, rev, e0, int(L * rng.uniform(1.0, 3.0))
            else:
                cc, rr = int(rng.choice(n_contigs, p=p_ctg)), int(rng.integers(0, 2))
                s0, chn = int(rng.integers(2, ctg_len[cc] - 2)), int(rng.integers(kmer, max(kmer + 1, L)))
            cand["target_offset"].append(ctg_off[cc]), cand["target_length"].append(ctg_len[cc])
            cand["query_offset"].append(q_off[-1]), cand["query_length"].append(L)
            cand["anchor_offset"].append(len(anch_block)), cand["anchor_count"].append(n_anchor)
            cand["reverse"].append(rr), cand["chn"].append(chn), cand["vid"].append(cc * 2 + rr)
            cand["vs"].append(s0 - 2)
            for t in range(n_anchor):
                k = kmer - 1 + (L - kmer) * t // max(1, n_anchor - 1)
                if j != true_at:
                    k = int(rng.integers(kmer - 1, L))
                pos = s0 + 3 * k if j == true_at else int(rng.integers(0, ctg_len[cc]))
                anch_block.append(max(0, min(pos, int(ctg_len[cc]) - 1)) // bw * bw), anch_qend.append(k)
        q_off.append(q_off[-1] + L)
        cand_off.append(len(cand["vid"]))
    out = {"genome": g, "queries": np.concatenate(queries).astype(np.uint8), "qlen": qlen,
           "cand_off": np.array(cand_off, dtype=np.int32), "anchor_block": np.array(anch_block, dtype=np.int64),
           "anchor_qend": np.array(anch_qend, dtype=np.int32)}
    for k2, v in cand.items():
        out[k2] = np.array(v, dtype=np.uint32 if k2 == "vid" else np.int64)
    return out


def genome_with_genes(n_prot=20000, genome_len=100_000_000, n_ctg=24, seed=21, gc=0.42, repeat_frac=0.05,
                      med_len=400, max_len=3000):
    """Genome-scale workload: proteins with log-normal lengths (median med_len, cap max_len) and a
    genome of genome_len bases (GC gc) in n_ctg contigs, with repeat_frac of it in low-complexity
    tandem repeats (units of 1-6 bases) and one planted gene model per protein on a random strand:
    Poisson(6) exons (at least 1), GT..AG introns log-normal (median 1 kb, cap 50 kb; scaled down
    together when the genes would not fit in 90% of the genome), 5-35% amino-acid substitutions.
    Returns (contig names, contig sequences as bytes, protein names, protein sequences as bytes,
    intron scale)."""
    rng = np.random.default_rng(seed)
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    g = NT[rng.choice(4, size=genome_len, p=p)]
    # low-complexity tandem repeats
    n_rep, covered = 0, 0
    while covered < repeat_frac * genome_len:
        unit = NT[rng.integers(0, 4, size=int(rng.integers(1, 7)))]
        L = int(rng.integers(50, 2000))
        s = int(rng.integers(0, genome_len - L))
        g[s:s + L] = np.resize(unit, L)
        covered += L
        n_rep += 1
    lens = np.minimum(max_len, np.maximum(30, np.round(np.exp(rng.normal(np.log(med_len), 0.6, size=n_prot))))).astype(int)
    prots = [rand_protein(rng, int(L)) for L in lens]
    n_exon = np.maximum(1, rng.poisson(6, size=n_prot))
    introns = [np.minimum(50_000, np.maximum(40, np.exp(rng.normal(np.log(1000), 0.8, size=k - 1)))) for k in n_exon]
    cds_total = int(3 * lens.sum())
    intr_total = float(sum(x.sum() for x in introns))
    budget = 0.9 * genome_len - cds_total - 200 * n_prot
    scale = min(1.0, budget / intr_total) if intr_total > 0 else 1.0
    genes = []
    for k in range(n_prot):
        rate = rng.uniform(0.05, 0.35)
        cds = reverse_translate(rng, substitute(rng, prots[k], rate))
        ni = len(introns[k])
        cuts = sorted(rng.choice(np.arange(1, len(cds)), size=min(ni, len(cds) - 1), replace=False).tolist()) if ni else []
        seq, last = bytearray(), 0
        for c, il in zip(cuts, introns[k]):
            seq += cds[last:c]
            L = max(8, int(il * scale))
            seq += b"GT" + rand_dna(rng, L - 4) + b"AG"
            last = c
        seq += cds[last:]
        if rng.random() < 0.5:   # reverse strand
            seq = bytes(seq[::-1]).translate(bytes.maketrans(b"ACGT", b"TGCA"))
        genes.append(bytes(seq))
    # place the genes one after another at random spacing, in random order
    order = rng.permutation(n_prot)
    free = genome_len - sum(len(x) for x in genes)
    gaps = rng.dirichlet(np.ones(n_prot + 1)) * free
    pos = 0.0
    for j, k in enumerate(order):
        pos += gaps[j]
        s = int(pos)
        g[s:s + len(genes[k])] = np.frombuffer(genes[k], dtype=np.uint8)
        pos += len(genes[k])
    cuts = np.sort(rng.choice(np.arange(10_000, genome_len - 10_000), n_ctg - 1, replace=False))
    bounds = np.concatenate([[0], cuts, [genome_len]])
    gb = g.tobytes()
    ctg_names = ["chr%d" % (i + 1) for i in range(n_ctg)]
    ctgs = [gb[bounds[i]:bounds[i + 1]] for i in range(n_ctg)]
    names = ["p%05d" % i for i in range(n_prot)]
    return ctg_names, ctgs, names, prots, scale


def write_fasta(path, names, seqs, width=0):
    """FASTA (one line per sequence unless width > 0); gzip-compressed when the path ends in .gz"""
    import gzip
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "wb") as f:
        for n, s in zip(names, seqs):
            f.write(b">" + n.encode() + b"\n")
            if width:
                for i in range(0, len(s), width):
                    f.write(s[i:i + width] + b"\n")
            else:
                f.write(s + b"\n")
