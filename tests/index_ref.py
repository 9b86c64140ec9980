"""CPU restatement (test oracle) of the k-mer block index: mp_idx_boff, mp_sketch_nt4 with
mp_sketch_clean_orf and mp_hash32_mask, build_bidx (index.c:14-29, 52-97; sketch.c:7-16,
108-169). Plain loops over the ORFs of pssm_ref.orf_walk (the same frame walk), for small
genomes; pinned against the compiled reference (oracle/_ref, tests/golden/index_cases.json.gz)
in tests/test_index_cpu.py."""
import numpy as np

import pssm_ref

M32 = 0xffffffff


def hash32_mask(key, mask):
    key = (key + (~(key << 15) & M32)) & mask
    key ^= key >> 10
    key = (key + (key << 3)) & mask
    key ^= key >> 6
    key = (key + (~(key << 11) & M32)) & mask
    key ^= key >> 16
    return key


def boff(ctg_len, bbit):
    bo, b = [], 0
    for ln in ctg_len:
        nb = (ln + (1 << bbit) - 1) >> bbit
        bo += [b, b + nb]
        b += 2 * nb
    return bo + [b]


def oriented(genome, off, ln, rev):
    s = [int(x) for x in genome[off:off + ln]]
    return [c if c >= 4 else 3 - c for c in reversed(s)] if rev else s


def sketch_nt4(seq, min_aa_len, kmer, mod_bit, bbit, bo, codon, codon13):
    """the sorted unique entries (hash >> mod_bit) << 32 | block of one oriented strand"""
    mask_k, mask_mod = (1 << 4 * kmer) - 1, (1 << mod_bit) - 1
    orfs, _ = pssm_ref.orf_walk(seq, min_aa_len, codon)
    out = set()
    for st, k in orfs:
        x = 0
        for t in range(k):
            i = st + 3 * t
            x = ((x << 4) | codon13[seq[i] << 4 | seq[i + 1] << 2 | seq[i + 2]]) & mask_k
            if t + 1 >= kmer:
                y = hash32_mask(x, mask_k)
                if y & mask_mod == 0:
                    out.add((y >> mod_bit) << 32 | (((i + 2) >> bbit) + bo))
    return sorted(out)


def index_build(genome, ctg_off, ctg_len, min_aa_len, kmer, mod_bit, bbit, codon, codon13):
    """(ki, kb, bo, keys): bucket starts, entries, block offsets, and the sorted unique entries"""
    bo = boff(ctg_len, bbit)
    keys = []
    for j in range(2 * len(ctg_len)):
        seq = oriented(genome, ctg_off[j >> 1], ctg_len[j >> 1], j & 1)
        keys += sketch_nt4(seq, min_aa_len, kmer, mod_bit, bbit, bo[j], codon, codon13)
    keys.sort()   # a bucket holds its strands in order; block numbers grow with the strand
    n_bucket = 1 << (4 * kmer - mod_bit)
    cnt = np.bincount(np.array([k >> 32 for k in keys], dtype=np.int64), minlength=n_bucket)
    ki = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    kb = [k & M32 for k in keys]
    return ki, kb, bo, keys


def sketch_prot(seq, kmer, mod_bit, aa13):
    """mp_sketch_prot (sketch.c:18-38): (hash >> mod_bit) << 32 | position of the kept k-mers"""
    mask_k, mask_mod = (1 << 4 * kmer) - 1, (1 << mod_bit) - 1
    out, x, l = [], 0, 0
    for i, ch in enumerate(seq):
        c = aa13[ch]
        if c < 14:
            x = ((x << 4) | c) & mask_k
            l += 1
            if l >= kmer:
                y = hash32_mask(x, mask_k)
                if y & mask_mod == 0:
                    out.append((y >> mod_bit) << 32 | i)
        else:
            x = l = 0
    return sorted(out)


def anchors(ki, kb, seq, kmer, mod_bit, max_occ, aa13):
    """the sorted anchors of one query (ASCII bytes), as map.c:583-597 collects them with the
    default options: the sketch, the boxplot repeat cutoff (mp_cal_max_occ, map.c:304-318) capped at
    max_occ, every bucket entry of the seeds under it"""
    sd = sketch_prot(seq, kmer, mod_bit, aa13)
    n_kb = len(kb)

    def size(b):
        return (ki[b + 1] if b + 1 < len(ki) else n_kb) - ki[b]
    occ = max_occ
    if len(sd) >= 8:
        cnt = sorted(size(s >> 32) for s in sd)
        q25, q75 = cnt[int(len(sd) * .25 + .499)], cnt[int(len(sd) * .75 + .499)]
        occ = min(int(q75 + (q75 - q25) * 1.5 + 10.), max_occ)
    out = []
    for s in sd:
        b, pos = s >> 32, s & M32
        if size(b) <= occ:
            out += [int(kb[ki[b] + j]) << 32 | pos for j in range(size(b))]
    return sorted(out)
