"""CPU restatements (test oracle) of the PSSM diagonal evidence and the six-frame ORF walk.

diag_seg / chain_pssm_ev follow mp_diag_seg and mp_chain_pssm_ev (map.c:91-139); orf_walk follows
ps_orf_worker and ps_add_orf (prescreen.c:87-132). Plain loops, for small inputs; pinned against
the compiled reference (oracle/_ref, tests/golden/pssm_chain_cases.json.gz and orf_cases.json.gz)
and by closed forms in tests/test_oracle_cpu.py.
"""


def _codon_aa(nt, p, codon, aa_x):
    b0, b1, b2 = nt[p], nt[p + 1], nt[p + 2]
    return aa_x if b0 > 3 or b1 > 3 or b2 > 3 else codon[b0 << 4 | b1 << 2 | b2]


def diag_seg(nt, pssm, qlen, kmer, x, q, xdrop, codon, aa_x):
    """(score, aa_end) of the two-sided X-drop segment through one anchor (map.c:97-120);
    pssm is a flat list of qlen x 22 values"""
    l_nt = len(nt)
    p0, q0 = x + 1 - 3 * kmer, q + 1 - kmer
    sc = best = 0
    best_j = q0 - 1
    for j in range(q0, qlen):
        p = p0 + 3 * (j - q0)
        if p < 0:
            continue
        if p + 2 >= l_nt:
            break
        sc += pssm[j * 22 + _codon_aa(nt, p, codon, aa_x)]
        if sc > best:
            best, best_j = sc, j
        elif best - sc > xdrop:
            break
    tot, aa_e = best, best_j + 1
    sc = best = 0
    for j in range(q0 - 1, -1, -1):
        p = p0 + 3 * (j - q0)
        if p < 0:
            break
        if p + 2 >= l_nt:
            continue
        sc += pssm[j * 22 + _codon_aa(nt, p, codon, aa_x)]
        if sc > best:
            best = sc
        elif best - sc > xdrop:
            break
    return tot + best, aa_e


def chain_pssm_ev(nt, pssm, qlen, kmer, anchors, min_seg, xdrop, codon, aa_x):
    """(evidence, [(score, aa_end) or None per anchor]) of mp_chain_pssm_ev (map.c:123-139)"""
    segs, ev, per = [], 0, []
    for av in anchors:
        x, q = av >> 32, av & 0xffffffff
        q0 = q + 1 - kmer
        d = (x + 1 - 3 * kmer) - 3 * q0
        if any(sd == d and q0 <= se for sd, se in segs):
            per.append(None)
            continue
        sc, e = diag_seg(nt, pssm, qlen, kmer, x, q, xdrop, codon, aa_x)
        per.append((sc, e))
        if len(segs) < 64:
            segs.append((d, e))
        if sc >= min_seg:
            ev += sc
    return ev, per


def orf_walk(seq, min_orf, codon):
    """[(start, codons)], residues: the ORFs of one oriented sequence (codes 0..4) in emission order"""
    orfs, aa = [], []

    def emit(e, k):
        st = e + 1 - 3 * k
        orfs.append((st, k))
        for t in range(k):
            c = seq[st + 3 * t:st + 3 * t + 3]
            aa.append(codon[c[0] << 4 | c[1] << 2 | c[2]])

    e, k = [-1] * 3, [0] * 3
    p, cod, l = 1, 0, 0
    for i, b in enumerate(seq):
        if p == 3:
            p = 0
        if b < 4:
            cod = (cod << 2 | b) & 0x3f
            l += 1
            if l >= 3:
                if codon[cod] >= 20:
                    if k[p] >= min_orf:
                        emit(e[p], k[p])
                    k[p], e[p] = 0, -1
                else:
                    e[p] = i
                    k[p] += 1
        else:
            for f in range(3):
                if k[f] >= min_orf:
                    emit(e[f], k[f])
                k[f], e[f] = 0, -1
            l, cod = 0, 0
        p += 1
    for f in range(3):
        if k[f] >= min_orf:
            emit(e[f], k[f])
    return orfs, aa


def ps_gapped(pssm, L, orf, diag, band, go, ge):
    """banded local affine score of a profile (flat L x 22) against an ORF (prescreen.c:346-382)"""
    NEG = -0x3fffffff
    n = len(orf)
    H, F = [0] * (n + 1), [NEG] * (n + 1)
    best = 0
    for i in range(L):
        lo = i + diag - band if band > 0 else 0
        hi = i + diag + band if band > 0 else n - 1
        lo, hi = max(lo, 0), min(hi, n - 1)
        if lo > hi:
            continue
        e = NEG
        h_diag = H[lo - 1] if lo > 0 else 0
        for t in range(lo, hi + 1):
            h_up = H[t]
            h = h_diag + pssm[i * 22 + orf[t]]
            f = max(F[t] - ge, h_up - go)
            F[t] = f
            h = max(h, f, e, 0)
            best = max(best, h)
            e = max(e - ge, h - go)
            h_diag = h_up
            H[t] = h
    return best
