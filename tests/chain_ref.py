"""CPU restatement (test oracle) of the anchor chaining mp_chain (chain.c:7-219): the score of
joining two anchors (comput_sc, with mp_log2 of mppriv.h:143-151 in float32), the chaining DP
with its skip heuristics, the greedy backtrack (mp_chain_backtrack, mp_chain_bk_end) and the
final ordering of chains by target position (compact_a), with the reference's own (unstable)
radix sort for the two orderings that can tie. Plain loops, for small inputs; pinned
against the compiled reference (oracle/_ref, tests/golden/anchor_chain_cases.json.gz) in
tests/test_chain_cpu.py."""
import struct

import numpy as np

INT32_MIN = -2 ** 31
INT32_MAX = 2 ** 31 - 1
BLOCK_BONUS = 2   # MP_BLOCK_BONUS (protlocus.h:47)
f32 = np.float32


def _i32(x):
    return (x + 2 ** 31) % 2 ** 32 - 2 ** 31


def log2f(x):
    """mp_log2 of a float32 x >= 2: exponent plus a quadratic in the mantissa"""
    bits = struct.unpack("<I", struct.pack("<f", float(f32(x))))[0]
    lg = f32(((bits >> 23) & 255) - 128)
    m = struct.unpack("<f", struct.pack("<I", (bits & ~(255 << 23)) + (127 << 23)))[0]
    m = f32(m)
    return f32(lg + f32(f32(f32(f32(-0.34484843) * m) + f32(2.02466578)) * m) - f32(0.67487759))


def comput_sc(ai, aj, max_dist_x, max_dist_y, bw, coef, is_spliced, bbit, kmer):
    dq = _i32((ai & 0xffffffff) - (aj & 0xffffffff))
    dq3 = _i32(dq * 3)
    if dq <= 0 or dq3 > max_dist_x or dq > max_dist_y:
        return INT32_MIN
    dd_signed = 0
    if bbit > 0:
        bs = 1 << bbit
        dr3 = _i32(((ai >> 32) - (aj >> 32)) << bbit)
        if dr3 - bs <= dq3 <= dr3 + bs:
            dd = 0
        elif dq3 < dr3 - bs:
            dd = dr3 - bs - dq3
            dd_signed = -dd
        else:
            dd = dq3 - (dr3 + bs)
            dd_signed = dd
    else:
        dr3 = _i32((ai >> 32) - (aj >> 32))
        if dr3 == 0:
            return INT32_MIN
        dd = abs(dr3 - dq3)
        dd_signed = dq3 - dr3
    if dd > bw:
        return INT32_MIN
    if bbit > 0:
        sc = min(kmer, dq)
    elif kmer <= dq and kmer * 3 <= dr3:
        sc = kmer
    else:
        dr = int(dr3 / 3)   # C division truncates toward zero
        q = dr3 - dr * 3
        sc = min(min(dr, dq), kmer)
        if q != 0:
            sc -= 1
    if dd > 0:
        lin = f32(f32(dd) * f32(0.33334))
        lg = f32(f32(coef) * f32(log2f(f32(dd + 1)) - f32(1.0))) + f32(1.0) if dd >= 2 else f32(dd)
        lg = f32(lg)
        if is_spliced and dd_signed < 0:
            sc -= int(min(lin, lg))
        else:
            sc -= int(f32(lin + lg))
    if bbit > 0 and (ai >> 32) == (aj >> 32) and dd == 0:
        sc += BLOCK_BONUS
    return sc


RS_MIN_SIZE, RS_MAX_BITS = 64, 8   # ksort.h:109-110


def _insertsort(arr, beg, end):
    for i in range(beg + 1, end):
        if arr[i][0] < arr[i - 1][0]:
            tmp, j = arr[i], i
            while j > beg and tmp[0] < arr[j - 1][0]:
                arr[j] = arr[j - 1]
                j -= 1
            arr[j] = tmp


def _rs_sort(arr, beg, end, n_bits, s):
    size = 1 << n_bits
    m = size - 1
    b, e = [beg] * size, [beg] * size
    for i in range(beg, end):
        e[(arr[i][0] >> s) & m] += 1
    for k in range(1, size):
        e[k] += e[k - 1] - beg
        b[k] = e[k - 1]
    k = 0
    while k != size:   # cycle-leader permutation into the buckets, in place (not stable)
        if b[k] != e[k]:
            lk = (arr[b[k]][0] >> s) & m
            if lk != k:
                tmp = arr[b[k]]
                while True:
                    swap, tmp = tmp, arr[b[lk]]
                    arr[b[lk]] = swap
                    b[lk] += 1
                    lk = (tmp[0] >> s) & m
                    if lk == k:
                        break
                arr[b[k]] = tmp
                b[k] += 1
            else:
                b[k] += 1
        else:
            k += 1
    b[0] = beg
    for k in range(1, size):
        b[k] = e[k - 1]
    if s:
        s = s - n_bits if s > n_bits else 0
        for k in range(size):
            if e[k] - b[k] > RS_MIN_SIZE:
                _rs_sort(arr, b[k], e[k], n_bits, s)
            elif e[k] - b[k] > 1:
                _insertsort(arr, b[k], e[k])


def radix_sort_128x(arr):
    """radix_sort_mp128x (ksort.h:112-162 on mp128_t by .x): in-place MSD radix sort of (x, y)
    pairs by x alone, with insertion sort for short buckets; the order of equal x follows the
    algorithm, so it is restated as it is"""
    if len(arr) <= RS_MIN_SIZE:
        _insertsort(arr, 0, len(arr))
    else:
        _rs_sort(arr, 0, len(arr), RS_MAX_BITS, 7 * RS_MAX_BITS)


def chain(a, max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, coef, is_spliced, kmer, bbit):
    """(u, anchors) of mp_chain: u = score << 32 | count per chain, in the returned order"""
    n = len(a)
    if n == 0:
        return [], []
    max_dist_x = max(max_dist_x, bw)
    if max_dist_y < bw and not is_spliced:
        max_dist_y = bw
    max_drop = INT32_MAX if is_spliced else bw
    f, p, v, t = [0] * n, [0] * n, [0] * n, [0] * n
    st, hi, hf = 0, -1, 0
    for i in range(n):
        max_j, max_f, n_skip = -1, kmer, 0
        while st < i and ((a[i] >> 32) - (a[st] >> 32)) << bbit > max_dist_x:
            st += 1
        if hi >= 0 and hi >= st:
            sc = hf + comput_sc(a[i], a[hi], max_dist_x, max_dist_y, bw, coef, is_spliced, bbit, kmer)
            sc = _i32(sc)
            if sc > max_f:
                max_f, max_j = sc, hi
        else:
            hf, hi = 0, -1
        if i - st > max_iter:
            st = i - max_iter
        for j in range(i - 1, st - 1, -1):
            sc = comput_sc(a[i], a[j], max_dist_x, max_dist_y, bw, coef, is_spliced, bbit, kmer)
            if sc == INT32_MIN:
                continue
            sc = _i32(sc + f[j])
            if sc > max_f:
                max_f, max_j = sc, j
                if n_skip > 0:
                    n_skip -= 1
            elif t[j] == i:
                n_skip += 1
                if n_skip > max_skip:
                    break
            if p[j] >= 0:
                t[p[j]] = i
        f[i], p[i] = max_f, max_j
        v[i] = v[max_j] if max_j >= 0 and v[max_j] > max_f else max_f
        if hf < max_f:
            hf, hi = max_f, i
    # backtrack (mp_chain_backtrack): candidates sorted by score, best first (z[k].x holds the
    # int32 score widened to uint64, so a negative score sorts last)
    z = [(f[i] % 2 ** 64, i) for i in range(n) if f[i] >= min_sc]
    radix_sort_128x(z)
    z = [(_i32(x), i) for x, i in z]
    if not z:
        return [], []

    def bk_end(k):
        i = z[k][1]
        end_i, max_i, max_s = -1, i, 0
        if i < 0 or t[i] != 0:
            return i
        while True:
            t[i] = 2
            end_i = i = p[i]
            s = z[k][0] if i < 0 else z[k][0] - f[i]
            if s > max_s:
                max_s, max_i = s, i
            elif max_s - s > max_drop:
                break
            if not (i >= 0 and t[i] == 0):
                break
        i = z[k][1]
        while i >= 0 and i != end_i:
            t[i] = 0
            i = p[i]
        return max_i
    t = [0] * n
    u, vv = [], []
    for k in range(len(z) - 1, -1, -1):
        if t[z[k][1]] != 0:
            continue
        n_v0 = len(vv)
        end_i = bk_end(k)
        i = z[k][1]
        while i != end_i:
            vv.append(i)
            t[i] = 1
            i = p[i]
        sc = z[k][0] if i < 0 else z[k][0] - f[i]
        if sc >= min_sc and len(vv) > n_v0 and len(vv) - n_v0 >= min_cnt:
            u.append(((sc & 0xffffffff) << 32) | (len(vv) - n_v0))
        else:
            del vv[n_v0:]
    if not u:
        return [], []
    # compact_a: each chain's anchors in increasing order, chains by their first target block
    b, k = [], 0
    for x in u:
        m = x & 0xffffffff
        b += [a[vv[k + m - 1 - j]] for j in range(m)]
        k += m
    w, k = [], 0
    for i, x in enumerate(u):
        w.append((b[k] >> 32, k << 32 | i))
        k += x & 0xffffffff
    radix_sort_128x(w)
    u2, a2 = [], []
    for _, y in w:
        k, i = y >> 32, y & 0xffffffff
        u2.append(u[i])
        a2 += b[k:k + (u[i] & 0xffffffff)]
    return u2, a2
