"""CPU restatement (test oracle) of the --gpu-prefilter candidate selection.

calibrate, adaptive_k and select_gapless follow mp_gapless_calibrate, mp_gapless_adaptive_k and
mp_select_gapless with its small-limit path (hit.c:42-263). select_batch drives them the way
mp_gpu_prefilter_batch and mp_gpu_prefilter_select do (map.c:1075-1083, 507-521). The code uses
plain loops, so keep the inputs small. It is pinned against the compiled reference
(oracle/_ref, tests/golden/gapless_select_cases.json.gz) in tests/test_gsel_cpu.py.
"""
import functools

INT32_MIN = -2 ** 31
TIE_PCT = 80          # MP_GAPLESS_TIE_PCT (mppriv.h:94)
DUP_RICH = 0.25       # MP_GAPLESS_DUP_RICH, a float constant (mppriv.h:95)
ADAPT_MAX_K = 4       # MP_GAPLESS_ADAPT_MAX_K (mppriv.h:96)


def _f32(x):
    import numpy as np
    return float(np.float32(x))


def rank_key(i, sc, chn, vid, vs):
    """mp_gapless_rank_cmp (hit.c:42-53): score desc, chain score desc, vid asc, vs asc, index asc"""
    return (-sc[i], -chn[i], vid[i], vs[i], i)


def calibrate(cand_off, scores):
    """(noise, n_query, dup_frac) of mp_gapless_calibrate (hit.c:149-189)"""
    n_query = len(cand_off) - 1
    bg, n_scored = [], 0

    def best2(q):
        best = second = INT32_MIN
        for i in range(cand_off[q], cand_off[q + 1]):
            if scores[i] > best:
                second, best = best, scores[i]
            elif scores[i] > second:
                second = scores[i]
        return best, second
    for q in range(n_query):
        if cand_off[q + 1] - cand_off[q] < 2:
            continue
        best, _ = best2(q)
        if best <= 0:
            continue
        n_scored += 1
        bg += [scores[i] for i in range(cand_off[q], cand_off[q + 1]) if 0 < scores[i] < best]
    noise = sorted(bg)[len(bg) // 2] if len(bg) >= 64 else 0
    n_conf = n_multi = 0
    for q in range(n_query):
        if cand_off[q + 1] - cand_off[q] < 2:
            continue
        best, second = best2(q)
        if best <= 0 or best < 4 * noise:
            continue
        n_conf += 1
        if second > 0 and second * 100 >= best * TIE_PCT:
            n_multi += 1
    dup = _f32(_f32(n_multi) / _f32(n_conf)) if n_conf > 0 else 0.0
    return noise, n_scored, dup


def adaptive_k(scores, qlen, max_k, noise, dup_frac):
    """mp_gapless_adaptive_k (hit.c:195-215)"""
    n = len(scores)
    if n <= 0 or max_k <= 0:
        return 0
    best = max(scores)
    if best <= 0 or best < 2 * noise or best * 4 < qlen:
        target = max_k
    else:
        thr = best * TIE_PCT // 100
        thr = max(thr, 2 * noise)
        target = sum(1 for s in scores if s >= thr)
        if dup_frac >= _f32(DUP_RICH):
            target += 1
    return max(1, min(target, max_k, n))


def _chain_guard(sc, chn):
    """mp_gapless_chain_guard (hit.c:71-85)"""
    best = top = -1
    for i in range(len(sc)):
        if best < 0 or chn[i] > chn[best] or (chn[i] == chn[best] and sc[i] > sc[best]):
            best = i
        if top < 0 or sc[i] > sc[top] or (sc[i] == sc[top] and chn[i] > chn[top]):
            top = i
    if best < 0 or best == top or chn[top] * 10 >= chn[best] * 8:
        return -1
    return best


def select_gapless(top_k, min_per_vid, sc, chn, vid, vs):
    """the indices mp_select_gapless keeps (hit.c:217-263 and its small path, 87-136)"""
    n = len(sc)
    if n <= 0:
        return []
    limit = min(n, top_k)
    guard = _chain_guard(sc, chn) if limit < n else -1
    key = functools.partial(rank_key, sc=sc, chn=chn, vid=vid, vs=vs)
    sel = [False] * n
    if limit <= 8:      # streaming top-`limit` with the per-vid floor, then the fill
        top = []
        if min_per_vid > 0:
            for i in range(n):
                same = [t for t in top if vid[t] == vid[i]]
                if len(same) >= min_per_vid:
                    worst = max(same, key=key)
                    if key(i) >= key(worst):
                        continue
                    top.remove(worst)
                top.append(i)
                top.sort(key=key)
                del top[limit:]
            for t in top:
                sel[t] = True
        if len(top) < limit:
            fill = sorted((i for i in range(n) if not sel[i]), key=key)[:limit - len(top)]
            for t in fill:
                sel[t] = True
    else:
        order = sorted(range(n), key=key)
        kept, groups = 0, {}
        for i in order:
            if kept >= limit or min_per_vid <= 0:
                break
            if groups.get(vid[i], 0) < min_per_vid:
                sel[i] = True
                groups[vid[i]] = groups.get(vid[i], 0) + 1
                kept += 1
        for i in order:
            if kept >= limit:
                break
            if not sel[i]:
                sel[i] = True
                kept += 1
    if guard >= 0:
        sel[guard] = True
    return [i for i in range(n) if sel[i]]


def select_batch(cand_off, qlen, scores, chn, vid, vs, top_k, min_per_vid, adaptive):
    """(keep flags, K per query, (noise, n_query, dup_frac)) for a mini-batch"""
    cal = calibrate(cand_off, scores) if adaptive else (0, 0, 0.0)
    keep = [0] * cand_off[-1]
    ks = []
    for q in range(len(cand_off) - 1):
        a, b = cand_off[q], cand_off[q + 1]
        if a == b:
            ks.append(0)
            continue
        k = top_k
        if adaptive:
            k = adaptive_k(scores[a:b], qlen[q], min(k, ADAPT_MAX_K), cal[0], cal[2])
        ks.append(k)
        for i in select_gapless(k, min_per_vid, scores[a:b], chn[a:b], vid[a:b], vs[a:b]):
            keep[a + i] = 1
    return keep, ks, cal
