"""CPU restatement of the 64 prescreen features (the definitions in features.hip), one
candidate at a time, for checking the GPU feature kernel on small inputs."""
import math

import numpy as np


def features(prot, dna, nt4, aa20, codon, mat):
    """prot: protein bytes (ASCII), dna: DNA bytes (ASCII or 0..4); tables and the 22 x 22 matrix
    ([target residue][protein residue]) as flat lists"""
    p = [aa20[b] for b in prot]
    w = [nt4[b] for b in dna]
    al, nl = len(p), len(w)
    first = {}
    for i in range(al - 2):
        x, y, z = p[i], p[i + 1], p[i + 2]
        if x < 20 and y < 20 and z < 20:
            first.setdefault((x * 20 + y) * 20 + z, i)
    ncod = nl // 3
    ncodf = [max(0, (nl - f) // 3) for f in range(3)]

    def aa_at(f, j):
        if j >= ncodf[f]:
            return 21
        o = f + 3 * j
        b0, b1, b2 = w[o], w[o + 1], w[o + 2]
        return codon[b0 << 4 | b1 << 2 | b2] if b0 < 4 and b1 < 4 and b2 < 4 else 21

    anch = [0, 0, 0]
    stops = [0, 0, 0]
    xcod = 0
    qhit, thit = set(), set()
    quart = [0, 0, 0, 0]
    hist = {}
    ds = ds2 = 0
    for f in range(3):
        for j in range(ncodf[f]):
            r0, r1, r2 = aa_at(f, j), aa_at(f, j + 1), aa_at(f, j + 2)
            if r0 == 20:
                stops[f] += 1
            if r0 == 21:
                xcod += 1
            if r0 < 20 and r1 < 20 and r2 < 20:
                q = first.get((r0 * 20 + r1) * 20 + r2)
                if q is not None:
                    anch[f] += 1
                    qhit.add(q)
                    thit.add(j)
                    quart[(4 * q) // al] += 1
                    d = j - q
                    key = (f, (d + al) >> 3)
                    hist[key] = hist.get(key, 0) + 1
                    ds += d
                    ds2 += d * d
    gc = sum(1 for b in w if b in (1, 2))
    gt = sum(1 for i in range(nl - 1) if w[i] == 2 and w[i + 1] == 3)
    ag = sum(1 for i in range(nl - 1) if w[i] == 0 and w[i + 1] == 2)
    nanch = sum(anch)
    dom = 0 if anch[0] >= anch[1] and anch[0] >= anch[2] else (1 if anch[1] >= anch[2] else 2)
    domc = [0] * 20
    for j in range(ncodf[dom]):
        r = aa_at(dom, j)
        if r < 20:
            domc[r] += 1
    pc = [0] * 20
    for r in p:
        if r < 20:
            pc[r] += 1
    # the four fullest diagonal bins (ties: smaller frame * nb + bin), scored by the best ungapped
    # segment over their 8 diagonals
    nb = (al + ncod) // 8 + 2
    counts = {}
    for (f, b), v in hist.items():
        counts[f * nb + b] = v
    seg = []
    for _ in range(4):
        if not counts or max(counts.values()) <= 0:
            seg.append(0)
            continue
        top = max(counts.values())
        ix = min(i for i, v in counts.items() if v == top)
        del counts[ix]
        f, b = ix // nb, ix % nb
        best = 0
        for dd in range(8):
            d = b * 8 - al + dd
            sc = 0
            for q in range(max(0, -d), min(al, ncodf[f] - d)):
                sc = max(0, sc + mat[aa_at(f, q + d) * 22 + p[q]])
                best = max(best, sc)
        seg.append(best)
    seg.sort(reverse=True)
    f32 = np.float32
    o = np.zeros(64, dtype=np.float32)
    fal, fnl = f32(al), f32(nl)
    o[0] = math.log2(1 + al)
    o[1] = math.log2(1 + nl)
    o[2] = fnl / (f32(3) * fal)
    for f in range(3):
        o[3 + f] = f32(anch[f]) / fal
    o[6] = f32(nanch) / fal
    for f in range(3):
        o[7 + f] = f32(stops[f]) / f32(max(1, ncodf[f]))
    o[10] = f32(xcod) / f32(max(1, sum(ncodf)))
    den = fnl if nl > 0 else f32(1)
    o[11] = f32(gc) / den
    o[12] = f32(16) * f32(gt) / den
    o[13] = f32(16) * f32(ag) / den
    o[14] = f32(anch[dom]) / f32(max(1, nanch))
    o[15] = f32(max(hist.values()) if hist else 0) / fal
    o[16] = f32(sum(1 for v in hist.values() if v >= 3)) / f32(10)
    o[17] = f32(len(qhit)) / fal
    o[18] = f32(len(thit)) / f32(max(1, ncod))
    sd = 0.0
    if nanch:
        m, m2 = ds / nanch, ds2 / nanch
        sd = math.sqrt(m2 - m * m) if m2 - m * m > 0 else 0.0
    o[19] = sd / al
    for r in range(20):
        o[20 + r] = f32(pc[r]) / fal
    for r in range(4):
        o[40 + r] = f32(seg[r]) / fal
    o[44] = f32(sum(seg)) / fal
    o[45] = f32(sum(1 for v in seg if v >= 20)) / f32(4)
    for r in range(14):
        o[46 + r] = f32(domc[r]) / f32(max(1, ncodf[dom]))
    for k in range(4):
        o[60 + k] = f32(quart[k]) / fal
    return o
