// Anchor chaining (mp_chain, chain.c:7-219) for a batch of queries on the device.
//
// The reference chains each query's sorted anchors on one host thread:
//   * a DP over anchors, looking back within max_dist_x, with two heuristics that make it
//     order-dependent: a counter of non-improving predecessors already claimed by a later anchor
//     (max_skip) and the best anchor so far as an extra predecessor;
//   * a greedy backtrack from the best scores down (mp_chain_backtrack, mp_chain_bk_end);
//   * the chains ordered by their first target block (compact_a).
// Its two sorts of (key, index) pairs (radix_sort_mp128x, ksort.h:112-162) are in-place MSD radix
// sorts that do not keep the order of equal keys, and equal scores and equal blocks are common;
// the candidates' order, and so the chains, depend on the exact permutation.
//
// Every query's chaining keeps that procedure's order of decisions; a wavefront runs one query
// (the batch fills the GPU with queries, largest first) and spreads over its 64 lanes what does
// not depend on that order:
//   * the DP's look-back: lane L holds anchor i-1-L (a, f, p shifted one lane per anchor), so the
//     nearest 64 predecessors are scored in parallel; a prefix maximum gives each lane the best
//     score before it, and only the few lanes that change the skip counter are walked in order.
//     The "claimed by a later anchor" marks (t[j] == i) only live within one anchor's scan, so
//     they are a 64-bit lane mask, or per-anchor words for windows longer than 64;
//   * the radix sort: key ranges and digit counts are counted by all lanes, a level where every
//     key has the same digit (the permutation is then the identity) is skipped, and buckets of at
//     most 64 keys are placed by their stable rank (what the insertion sort yields); the
//     cycle-leader pass that fixes the order of equal keys runs on one lane;
//   * the greedy backtrack runs on one lane; building and ordering the chains is lane-parallel.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int kRsMin = 64, kRsBits = 8, kRsSize = 1 << kRsBits;
// the sorted keys (chain scores, which start at kmer >= 1 and only grow, and target blocks) are
// below 2^32, so with identity levels skipped the recursion holds at most 4 levels (shifts 24, 16,
// 8 and 0)
constexpr int kFrames = 4;
constexpr int kBlockBonus = 2;   // MP_BLOCK_BONUS (protlocus.h:47)

using ChainParams = nsh_chain::Params;

struct ChArgs {
	const uint64_t *a; const int64_t *off; int32_t n_q;
	const int64_t *cnt;                // per query anchor count, or null: the offsets' differences
	const int32_t *order;              // queries by decreasing anchor count
	ChainParams cp;
	int32_t *f, *p, *t, *vv;           // per anchor
	uint8_t *hs;                       // per anchor: another anchor chains to it
	uint64_t *zx, *zy;                 // per anchor: sort pairs
	uint64_t *b;                       // per anchor: the chains' anchors before the final order
	uint64_t *u_out, *a_out;           // per anchor slot: chains (score << 32 | count), their anchors
	int32_t *n_u;                      // per query
};

struct ChLds {                         // one wavefront's shared state
	int32_t e[kFrames][kRsSize];            // radix sort frames: bucket ends (a bucket starts where the last ends)
	union {
		int32_t cur[kRsSize];               // the distribution pass's bucket cursors
		uint64_t sk[64];                    // the keys of a bucket of at most 64
	};
	int32_t fb[kFrames], fk[kFrames], fs[kFrames], frec[kFrames];   // per frame: range start, next bucket,
	                                                                // the children's shift, recursion
	uint64_t cmask, orv;
	int32_t nu, nv;
};

#define DEV __device__ __forceinline__

template <int CTRL, int ROWS = 0xf>
DEV int dpp(int old, int src) { return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWS, 0xf, false); }

// inclusive scans over lanes 0..63: shifts 1-8 inside each row of 16, then rows 15 and 31 onwards
DEV int wave_sum_incl(int v)
{
	v += dpp<0x111>(0, v);
	v += dpp<0x112>(0, v);
	v += dpp<0x114>(0, v);
	v += dpp<0x118>(0, v);
	v += dpp<0x142, 0xa>(0, v);
	v += dpp<0x143, 0xc>(0, v);
	return v;
}
DEV int wave_max_incl(int v)
{
	v = max(v, dpp<0x111>(INT_MIN, v));
	v = max(v, dpp<0x112>(INT_MIN, v));
	v = max(v, dpp<0x114>(INT_MIN, v));
	v = max(v, dpp<0x118>(INT_MIN, v));
	v = max(v, dpp<0x142, 0xa>(INT_MIN, v));
	v = max(v, dpp<0x143, 0xc>(INT_MIN, v));
	return v;
}
// lane L takes lane L-1's value, lane 0 takes first
DEV int shr1(int v, int first) { return dpp<0x138>(first, v); }
DEV uint64_t shr1_64(uint64_t v, uint64_t first)
{
	const uint32_t lo = (uint32_t)shr1((int)(uint32_t)v, (int)(uint32_t)first);
	const uint32_t hi = (uint32_t)shr1((int)(uint32_t)(v >> 32), (int)(uint32_t)(first >> 32));
	return (uint64_t)hi << 32 | lo;
}
DEV uint64_t readlane_64(uint64_t v, int l)
{
	return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32 |
		(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
}
DEV void wave_sync_lds()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// mp_log2 (mppriv.h:143-151): exponent plus a quadratic in the mantissa, for x >= 2
__device__ float log2_approx(float x)
{
	uint32_t i = __float_as_uint(x);
	float lg = (float)((int)((i >> 23) & 255) - 128);
	i &= ~(255u << 23);
	i += 127u << 23;
	const float m = __uint_as_float(i);
	lg += (-0.34484843f * m + 2.02466578f) * m - 0.67487759f;
	return lg;
}

// comput_sc (chain.c:112-151)
__device__ int32_t comput_sc(uint64_t ai, uint64_t aj, const ChainParams &c)
{
	const int32_t dq = (int32_t)(uint32_t)ai - (int32_t)(uint32_t)aj, dq3 = dq * 3;
	int32_t dr3, dd, sc, dd_signed = 0;
	if (dq <= 0 || dq3 > c.max_dist_x) return INT32_MIN;
	if (dq > c.max_dist_y) return INT32_MIN;
	if (c.bbit > 0) {
		const int32_t bs = 1 << c.bbit;
		dr3 = (int32_t)(((ai >> 32) - (aj >> 32)) << c.bbit);
		if (dq3 >= dr3 - bs && dq3 <= dr3 + bs) dd = dd_signed = 0;
		else if (dq3 < dr3 - bs) dd = dr3 - bs - dq3, dd_signed = -dd;
		else dd = dq3 - (dr3 + bs), dd_signed = dd;
	} else {
		dr3 = (int32_t)((ai >> 32) - (aj >> 32));
		if (dr3 == 0) return INT32_MIN;
		dd = dr3 > dq3? dr3 - dq3 : dq3 - dr3;
		dd_signed = dq3 - dr3;
	}
	if (dd > c.bw) return INT32_MIN;
	if (c.bbit > 0) {
		sc = c.kmer < dq? c.kmer : dq;
	} else if (c.kmer <= dq && c.kmer * 3 <= dr3) {
		sc = c.kmer;
	} else {
		const int32_t dr = dr3 / 3, q = dr3 - dr * 3;
		const int32_t dg = dr < dq? dr : dq;
		sc = dg < c.kmer? dg : c.kmer;
		if (q != 0) --sc;
	}
	if (dd > 0) {
		const float lin = (float)dd * .33334f;
		const float lg = dd >= 2? c.coef * (log2_approx((float)(dd + 1)) - 1.0f) + 1.0f : (float)dd;
		if (c.is_spliced && dd_signed < 0) sc -= (int)(lin < lg? lin : lg);
		else sc -= (int)(lin + lg);
	}
	if (c.bbit > 0 && ai >> 32 == aj >> 32 && dd == 0) sc += kBlockBonus;
	return sc;
}

// rs_insertsort (ksort.h:116-126) of at most 64 pairs: insertion by a strict comparison is a
// stable sort, so each pair goes to its stable rank
DEV void small_sort(uint64_t *x, uint64_t *y, int32_t beg, int32_t end, ChLds &L)
{
	const int lane = threadIdx.x, n = end - beg;
	uint64_t kx = 0, ky = 0;
	if (lane < n) kx = x[beg + lane], ky = y[beg + lane];
	L.sk[lane] = kx;
	wave_sync_lds();
	int r = 0;
	if (lane < n)
		for (int m = 0; m < n; ++m) {
			const uint64_t o = L.sk[m];
			r += o < kx || (o == kx && m < lane);
		}
	if (lane < n) x[beg + r] = kx, y[beg + r] = ky;
	__syncthreads();
}

// one rs_sort level (ksort.h:127-150) over [beg, end) into frame d: the key range first, so that
// levels where every key has the same digit (an identity permutation, after which the reference
// recurses into the same range) are skipped; then digit counts, bucket bounds, and the in-place
// cycle-leader permutation on lane 0, which fixes the order of equal keys
DEV bool distribute(uint64_t *x, uint64_t *y, int32_t beg, int32_t end, int32_t s, int d, ChLds &L)
{
	const int lane = threadIdx.x;
	const uint64_t x0 = x[beg];
	uint64_t df = 0;
	for (int32_t i = beg + lane; i < end; i += 64) df |= x[i] ^ x0;
	if (lane == 0) L.orv = 0;
	wave_sync_lds();
	if (df) atomicOr((unsigned long long *)&L.orv, (unsigned long long)df);
	wave_sync_lds();
	df = L.orv;
	if (df == 0) {   // every key equal: the identity at every level
		if (lane == 0) L.frec[d] = 0, L.fk[d] = kRsSize;
		wave_sync_lds();
		return true;
	}
	const int s2 = (63 - __builtin_clzll(df)) & ~(kRsBits - 1);
	if (s2 > s) return false;   // keys beyond the frames' reach
	s = s2;
	int32_t *B = L.cur, *E = L.e[d];
	for (int k = lane; k < kRsSize; k += 64) E[k] = 0;
	wave_sync_lds();
	for (int32_t i = beg + lane; i < end; i += 64) atomicAdd(&E[(x[i] >> s) & (kRsSize - 1)], 1);
	wave_sync_lds();
	{
		const int c0 = E[4 * lane], c1 = E[4 * lane + 1], c2 = E[4 * lane + 2], c3 = E[4 * lane + 3];
		const int tot = c0 + c1 + c2 + c3, ex = beg + wave_sum_incl(tot) - tot;
		wave_sync_lds();
		B[4 * lane] = ex, B[4 * lane + 1] = ex + c0, B[4 * lane + 2] = ex + c0 + c1, B[4 * lane + 3] = ex + c0 + c1 + c2;
		E[4 * lane] = ex + c0, E[4 * lane + 1] = ex + c0 + c1, E[4 * lane + 2] = ex + c0 + c1 + c2, E[4 * lane + 3] = ex + tot;
	}
	__syncthreads();
	// the cycle-leader pass, bucket by bucket: a run of keys already in their bucket advances the
	// cursor at once (all lanes look at the next 64); a key out of place starts a cycle on lane 0
	const int32_t m = kRsSize - 1;
	for (int32_t k = 0; k < kRsSize;) {
		const int32_t bk = B[k], ek = E[k];
		if (bk == ek) { ++k; continue; }
		const bool in = bk + lane < ek;
		const uint64_t out = __ballot(in && (int32_t)((x[bk + lane] >> s) & m) != k);
		const int32_t run = out? __builtin_ctzll(out) : min(64, ek - bk);
		wave_sync_lds();
		if (run > 0) {
			if (lane == 0) B[k] = bk + run;
			wave_sync_lds();
			continue;
		}
		if (lane == 0) {
			int32_t l = (int32_t)((x[bk] >> s) & m);
			uint64_t tx = x[bk], ty = y[bk];
			do {
				const uint64_t sx = tx, sy = ty;
				tx = x[B[l]], ty = y[B[l]];
				x[B[l]] = sx, y[B[l]] = sy;
				++B[l];
				l = (int32_t)((tx >> s) & m);
			} while (l != k);
			x[bk] = tx, y[bk] = ty;
			B[k] = bk + 1;
		}
		__syncthreads();
	}
	__syncthreads();
	if (lane == 0) L.fb[d] = beg, L.fs[d] = s > kRsBits? s - kRsBits : 0, L.frec[d] = s != 0, L.fk[d] = 0;
	wave_sync_lds();
	return true;
}

// radix_sort_mp128x (ksort.h:158-162) of n (x, y) pairs by x; the recursion runs on LDS frames,
// visiting each frame's buckets of two or more keys in order. False (nothing sorted) when a key
// reaches 2^32.
DEV bool radix_sort_pairs(uint64_t *x, uint64_t *y, int32_t n, ChLds &L)
{
	const int lane = threadIdx.x;
	if (n <= kRsMin) { small_sort(x, y, 0, n, L); return true; }
	if (!distribute(x, y, 0, n, (kFrames - 1) * kRsBits, 0, L)) return false;
	int depth = 0;
	while (depth >= 0) {
		const int k0 = L.fk[depth];
		if (!L.frec[depth] || k0 >= kRsSize) { --depth; continue; }
		const int base = k0 & ~63, kk = base + lane;
		const int32_t bb = kk? L.e[depth][kk - 1] : L.fb[depth], sz = L.e[depth][kk] - bb;
		const uint64_t m = __ballot(sz > 1) & (~0ull << (k0 & 63));
		const int lb = m? __builtin_ctzll(m) : 64;
		const int32_t beg = m? __builtin_amdgcn_readlane(bb, lb) : 0, size = m? __builtin_amdgcn_readlane(sz, lb) : 0;
		wave_sync_lds();
		if (lane == 0) L.fk[depth] = base + lb + (m? 1 : 0);
		wave_sync_lds();
		if (!m) continue;
		if (size > kRsMin) {
			distribute(x, y, beg, beg + size, L.fs[depth], depth + 1, L);
			++depth;
		} else small_sort(x, y, beg, beg + size, L);
	}
	return true;
}

__global__ void __launch_bounds__(64) chain_kernel(ChArgs A)
{
	__shared__ ChLds L;
	const int lane = threadIdx.x;
	const int32_t q = A.order[blockIdx.x];
	const int64_t o = A.off[q], n = A.cnt? A.cnt[q] : A.off[q + 1] - o;
	const uint64_t *a = A.a + o;
	int32_t *f = A.f + o, *p = A.p + o, *t = A.t + o, *vv = A.vv + o;
	uint64_t *zx = A.zx + o, *zy = A.zy + o, *b = A.b + o, *uo = A.u_out + o, *ao = A.a_out + o;
	uint8_t *hs = A.hs + o;
	ChainParams c = A.cp;
	if (n == 0) { if (lane == 0) A.n_u[q] = 0; return; }
	if (c.max_dist_x < c.bw) c.max_dist_x = c.bw;
	if (c.max_dist_y < c.bw && !c.is_spliced) c.max_dist_y = c.bw;
	const int32_t max_drop = c.is_spliced? INT32_MAX : c.bw;
	const uint64_t mdx = (uint64_t)(int64_t)c.max_dist_x;
	for (int64_t i = lane; i < n; i += 64) t[i] = 0;
	__syncthreads();
	// the DP (chain.c:178-208); lane L holds anchor i-1-L
	int64_t st = 0, hi = -1;
	int32_t hf = 0;
	uint64_t wa = 0;
	int32_t wf = 0, wp = -1;
	// the anchors ahead, 64 per register: lane L holds a[i0 + L] of this block and of the next
	uint64_t ablk = lane < n? a[lane] : 0, anext = 64 + lane < n? a[64 + lane] : 0;
	for (int64_t i = 0; i < n; ++i) {
		if (__builtin_expect((i & 63) == 0 && i > 0, 0)) {
			ablk = anext;
			anext = i + 64 + lane < n? a[i + 64 + lane] : 0;
		}
		const uint64_t ai = readlane_64(ablk, (int)(i & 63));
		bool stop = false;
		while (st < i - 64) {
			if (((ai >> 32) - (a[st] >> 32)) << c.bbit > mdx) ++st;
			else { stop = true; break; }
		}
		if (!stop && st < i) {   // the window's far end lies among the lanes: far lanes are the highest
			const int64_t j = i - 1 - lane;
			const uint64_t fm = __ballot(j >= st && ((ai >> 32) - (wa >> 32)) << c.bbit > mdx);
			if (fm) st = i - __builtin_ctzll(fm);
		}
		int32_t max_f = c.kmer, n_skip = 0;
		int64_t max_j = -1;
		if (hi >= 0 && hi >= st) {
			const uint64_t ah = hi >= i - 64? readlane_64(wa, (int)(i - 1 - hi)) : a[hi];
			const int32_t sc = hf + comput_sc(ai, ah, c);
			if (sc > max_f) max_f = sc, max_j = hi;
		} else hf = 0, hi = -1;
		if (i - st > c.max_iter) st = i - c.max_iter;
		for (int64_t c0 = i - 1; c0 >= st; c0 -= 64) {   // c0: lane 0's predecessor
			const int64_t j = c0 - lane;
			const bool in = j >= st;
			uint64_t aj = wa;
			int32_t fj = wf, pj = wp;
			if (c0 != i - 1) {
				__syncthreads();
				aj = in? a[j] : 0, fj = in? f[j] : 0, pj = in? p[j] : -1;
			}
			const int32_t sc = in? comput_sc(ai, aj, c) : INT32_MIN;
			const bool valid = sc != INT32_MIN;
			const uint64_t vm = __ballot(valid);
			if (!vm) continue;
			const int32_t scf = sc + fj;
			if (c0 == i - 1 && !(vm & (vm - 1))) {
				// one scoring predecessor among the nearest 64: nothing before it claims it, so it
				// only improves or not; its claim matters only to the chunks further down
				const int l = __builtin_ctzll(vm);
				if (valid && pj >= st && pj < c0 - 63)
					__hip_atomic_store(&t[pj], (int32_t)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
				const int32_t sl = __builtin_amdgcn_readlane(scf, l);
				if (sl > max_f) {
					max_f = sl, max_j = c0 - l;
					if (n_skip > 0) --n_skip;
				}
				continue;
			}
			// marks t[p[j]] = i of the lanes before: inside this chunk a lane mask, further down words
			// (read by the next chunk, after its barrier); claims below st are never read
			const bool inc = valid && pj >= st && pj >= c0 - 63;
			if (valid && pj >= st && pj < c0 - 63)
				__hip_atomic_store(&t[pj], (int32_t)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
			uint64_t cm = 0;
			if (__ballot(inc)) {
				if (lane == 0) L.cmask = 0;
				wave_sync_lds();
				if (inc) atomicOr((unsigned long long *)&L.cmask, 1ull << (c0 - pj));
				wave_sync_lds();
				cm = L.cmask;
				wave_sync_lds();
			}
			bool tm = (cm >> lane) & 1;
			if (c0 != i - 1 && in && !tm)
				tm = __hip_atomic_load(&t[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == (int32_t)i;
			const int32_t before = shr1(wave_max_incl(valid? scf : INT_MIN), INT_MIN);
			const bool imp = valid && scf > (before > max_f? before : max_f);
			const uint64_t im = __ballot(imp);
			uint64_t ev = __ballot(imp || (valid && tm));
			int brk = 64;
			while (ev) {
				const int l = __builtin_ctzll(ev);
				ev &= ev - 1;
				if ((im >> l) & 1) { if (n_skip > 0) --n_skip; }
				else if (++n_skip > c.max_skip) { brk = l; break; }
			}
			const uint64_t imb = brk == 64? im : im & ((1ull << brk) - 1);
			if (imb) {
				const int l = 63 - __builtin_clzll(imb);
				max_f = __builtin_amdgcn_readlane(scf, l), max_j = c0 - l;
			}
			if (brk < 64) break;
		}
		if (lane == 0) f[i] = max_f, p[i] = (int32_t)max_j;
		if (hf < max_f) hf = max_f, hi = i;
		wa = shr1_64(wa, ai), wf = shr1(wf, max_f), wp = shr1(wp, (int32_t)max_j);
	}
	__syncthreads();
	// the backtrack (mp_chain_backtrack, chain.c:27-82): candidates by score, best first
	int32_t n_z = 0;
	for (int64_t i0 = 0; i0 < n; i0 += 64) {
		const int64_t i = i0 + lane;
		const int32_t fi = i < n? f[i] : 0;
		const bool ok = i < n && fi >= c.min_sc;
		const uint64_t m = __ballot(ok);
		if (ok) {
			const int r = n_z + __builtin_popcountll(m & ((1ull << lane) - 1));
			zx[r] = (uint64_t)(int64_t)fi, zy[r] = (uint64_t)i;
		}
		n_z += __builtin_popcountll(m);
	}
	__syncthreads();
	if (n_z == 0) { if (lane == 0) A.n_u[q] = 0; return; }
	if (!radix_sort_pairs(zx, zy, n_z, L)) { if (lane == 0) A.n_u[q] = -1; return; }   // a score beyond int32
	// anchors no other anchor chains to (hs = 0): started from one with no predecessor either, the
	// backtrack only marks it (its chain of one is dropped when min_cnt > 1 or its score is not
	// positive), and no later walk can reach it, so such candidates are passed over
	for (int64_t i = lane; i < n; i += 64) t[i] = 0, hs[i] = 0;
	__syncthreads();
	for (int64_t i = lane; i < n; i += 64) if (p[i] >= 0) hs[p[i]] = 1;
	__syncthreads();
	int32_t bu = 0, bv = 0;   // lane 0's chain and anchor counts
	for (int32_t k0 = n_z - 1; k0 >= 0; k0 -= 64) {
		// the next 64 candidates, read by the lanes; a candidate already marked (t only turns
		// from 0 to 1 between candidates) is passed over as the reference does
		__syncthreads();   // lane 0's marks of the chunk before, for the lanes' reads of t
		const int32_t k = k0 - lane;
		bool live = false;
		int32_t zkl = 0, zsl = 0;
		if (k >= 0) {
			zkl = (int32_t)zy[k], zsl = (int32_t)zx[k];
			live = (p[zkl] >= 0 || hs[zkl] || (c.min_cnt <= 1 && zsl > 0 && zsl >= c.min_sc)) && t[zkl] == 0;
		}
		uint64_t lm = __ballot(live);
		while (lm) {
			const int l = __builtin_ctzll(lm);
			lm &= lm - 1;
			const int32_t zk = __builtin_amdgcn_readlane(zkl, l), zs = __builtin_amdgcn_readlane(zsl, l);
			if (lane != 0) continue;
			if (t[zk] != 0) continue;
			// mp_chain_bk_end (chain.c:8-25)
			int32_t end_i = -1, max_i = zk, max_s = 0, ii = zk;
			do {
				t[ii] = 2;
				end_i = ii = p[ii];
				const int32_t s = ii < 0? zs : zs - f[ii];
				if (s > max_s) max_s = s, max_i = ii;
				else if (max_s - s > max_drop) break;
			} while (ii >= 0 && t[ii] == 0);
			for (ii = zk; ii >= 0 && ii != end_i; ii = p[ii]) t[ii] = 0;
			end_i = max_i;
			const int32_t n_v0 = bv;
			for (ii = zk; ii != end_i; ii = p[ii]) vv[bv++] = ii, t[ii] = 1;
			const int32_t sc = ii < 0? zs : zs - f[ii];
			if (sc >= c.min_sc && bv > n_v0 && bv - n_v0 >= c.min_cnt) {
				uo[bu++] = (uint64_t)(uint32_t)sc << 32 | (uint32_t)(bv - n_v0);
				for (int32_t x = n_v0; x < bv; ++x) b[x] = (uint64_t)(n_v0 + bv - 1);   // x -> its mirror
			} else bv = n_v0;
		}
	}
	if (lane == 0) L.nu = bu, L.nv = bv;
	__syncthreads();
	const int32_t n_u = L.nu, n_v = L.nv;
	if (lane == 0) A.n_u[q] = n_u;
	if (n_u == 0) return;
	// compact_a (chain.c:84-110): each chain's anchors in increasing order, then the chains by
	// their first target block
	for (int32_t x = lane; x < n_v; x += 64) b[x] = a[vv[(int32_t)b[x] - x]];
	__syncthreads();
	for (int32_t i0 = 0, k0 = 0; i0 < n_u; i0 += 64) {
		const int32_t i = i0 + lane;
		const int32_t cnt = i < n_u? (int32_t)(uint32_t)uo[i] : 0;
		const int32_t incl = wave_sum_incl(cnt), k = k0 + incl - cnt;
		if (i < n_u) zx[i] = b[k] >> 32, zy[i] = (uint64_t)k << 32 | (uint32_t)i;
		k0 += __builtin_amdgcn_readlane(incl, 63);
	}
	__syncthreads();
	radix_sort_pairs(zx, zy, n_u, L);   // target blocks: below 2^32
	for (int32_t i0 = 0, k0 = 0; i0 < n_u; i0 += 64) {
		const int32_t i = i0 + lane;
		const int32_t j = i < n_u? (int32_t)(uint32_t)zy[i] : 0;
		const uint64_t uj = i < n_u? uo[j] : 0;
		const int32_t cnt = (int32_t)(uint32_t)uj;
		const int32_t incl = wave_sum_incl(cnt), k = k0 + incl - cnt;
		if (i < n_u) {
			const int64_t from = (int64_t)(zy[i] >> 32);
			for (int32_t x = 0; x < cnt; ++x) ao[k + x] = b[from + x];
			zx[i] = uj;
		}
		k0 += __builtin_amdgcn_readlane(incl, 63);
	}
	__syncthreads();
	for (int32_t i = lane; i < n_u; i += 64) uo[i] = zx[i];
}

// per query: the anchors held by its chains u[off[q], off[q] + nu[q])
__global__ void chain_kept_kernel(int32_t n_q, const int64_t *off, const int32_t *nu, const uint64_t *u,
		int64_t *kept)
{
	const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n_q) return;
	int64_t m = 0;
	for (int32_t i = 0; i < nu[q]; ++i) m += (uint32_t)u[off[q] + i];
	kept[q] = m;
}

// each query's chains and chain anchors, packed: the chains first (uoff), then the anchors (coff)
__global__ void chain_pack_kernel(const int64_t *off, const int32_t *nu, const int64_t *kept, const int64_t *uoff,
		const int64_t *coff, const uint64_t *u, const uint64_t *ao, uint64_t *pk)
{
	const int32_t q = blockIdx.x;
	for (int64_t i = threadIdx.x; i < nu[q]; i += blockDim.x) pk[uoff[q] + i] = u[off[q] + i];
	for (int64_t i = threadIdx.x; i < kept[q]; i += blockDim.x) pk[coff[q] + i] = ao[off[q] + i];
}

} // namespace

#define CHOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("chain: %s: %s", #x, hipGetErrorString(e_)); goto done; } } while (0)

namespace {
// device scratch kept between calls and grown on demand (hipMalloc of gigabytes per call costs
// more than the chaining of a mini-batch); calls on one process are serialised by g_ws_mu
std::recursive_mutex g_ws_mu;
struct Grow {
	void *p = nullptr;
	size_t cap = 0;
	hipError_t need(size_t bytes)
	{
		if (bytes <= cap) return hipSuccess;
		if (p) (void)hipFree(p), p = nullptr, cap = 0;
		const size_t c = bytes + bytes / 4;
		const hipError_t e = hipMalloc(&p, c);
		if (e == hipSuccess) cap = c;
		else p = nullptr;
		return e;
	}
};
Grow g_scratch, g_order, g_io;
size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
} // namespace

int nsh_chain::run(hipStream_t s, int32_t n_q, const int64_t *d_off, const int64_t *d_cnt, const int64_t *h_cnt,
		const uint64_t *d_a, int64_t n_a, const Params &prm, int32_t *d_nu, uint64_t *d_u, uint64_t *d_ao)
{
	if (n_q == 0) return 1;
	std::lock_guard<std::recursive_mutex> lk(g_ws_mu);
	int ok = 0;
	const size_t na = (size_t)(n_a > 0? n_a : 1);
	const size_t o4 = al256(na * 4), o8 = al256(na * 8);   // f, p, t, vv; zx, zy, b; hs
	uint64_t *d_zx = nullptr, *d_zy = nullptr, *d_b = nullptr;
	int32_t *d_f = nullptr, *d_p = nullptr, *d_t = nullptr, *d_vv = nullptr, *d_ord = nullptr;
	uint8_t *d_hs = nullptr;
	std::vector<int32_t> order(n_q);
	ChArgs A;
	{
		CHOK(g_scratch.need(4 * o4 + 3 * o8 + al256(na)));
		CHOK(g_order.need((size_t)n_q * 4));
		char *w = (char*)g_scratch.p;
		d_f = (int32_t*)w, d_p = (int32_t*)(w + o4), d_t = (int32_t*)(w + 2 * o4), d_vv = (int32_t*)(w + 3 * o4);
		d_zx = (uint64_t*)(w + 4 * o4), d_zy = (uint64_t*)(w + 4 * o4 + o8), d_b = (uint64_t*)(w + 4 * o4 + 2 * o8);
		d_hs = (uint8_t*)(w + 4 * o4 + 3 * o8);
		d_ord = (int32_t*)g_order.p;
		for (int32_t q = 0; q < n_q; ++q) order[q] = q;   // largest queries first
		std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return h_cnt[x] > h_cnt[y]; });
		CHOK(hipMemcpyAsync(d_ord, order.data(), (size_t)n_q * 4, hipMemcpyHostToDevice, s));
		A.a = d_a, A.off = d_off, A.cnt = d_cnt, A.n_q = n_q, A.order = d_ord, A.cp = prm;
		A.f = d_f, A.p = d_p, A.t = d_t, A.vv = d_vv, A.hs = d_hs, A.zx = d_zx, A.zy = d_zy, A.b = d_b;
		A.u_out = d_u, A.a_out = d_ao, A.n_u = d_nu;
		hipLaunchKernelGGL(chain_kernel, dim3(n_q), dim3(64), 0, s, A);
		CHOK(hipGetLastError());
		CHOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		ok = 1;
		{
			int32_t bad = 0;
			std::vector<int32_t> nu(n_q);
			CHOK(hipMemcpy(nu.data(), d_nu, (size_t)n_q * 4, hipMemcpyDeviceToHost));
			for (int32_t q = 0; q < n_q && !bad; ++q) bad = nu[q] < 0;
			if (bad) { nsh_set_error("chain: a chain score overflows int32"); ok = 0; }
		}
	}
done:
	(void)hipStreamSynchronize(s);
	return ok;
}

extern "C" int64_t nsh_chain_batch(int32_t n_q, const int64_t *a_off, const uint64_t *anchors, int32_t max_dist_x,
		int32_t max_dist_y, int32_t bw, int32_t max_skip, int32_t max_iter, int32_t min_cnt, int32_t min_sc,
		float chn_coef_log, int32_t is_spliced, int32_t kmer, int32_t bbit, int32_t *n_u, uint64_t *u, uint64_t *a_out)
{
	if (n_q < 0 || (n_q > 0 && (!a_off || !n_u)) || bbit < 0 || bbit > 30 || kmer < 1) {
		nsh_set_error("chain: bad arguments");
		return -1;
	}
	if (n_q == 0) return 0;
	const int64_t n_a = a_off[n_q];
	if (a_off[0] != 0 || n_a < 0) { nsh_set_error("chain: a_off must start at 0"); return -1; }
	std::vector<int64_t> cnt(n_q);
	for (int32_t q = 0; q < n_q; ++q) {
		if (a_off[q + 1] < a_off[q] || a_off[q + 1] - a_off[q] > INT32_MAX / 2) {
			nsh_set_error("chain: query %d has a bad anchor range", q);
			return -1;
		}
		cnt[q] = a_off[q + 1] - a_off[q];
	}
	if (n_a > 0 && (!anchors || !u || !a_out)) { nsh_set_error("chain: bad arguments"); return -1; }
	if (!nsh_ensure_device()) return -1;
	int64_t total = -1;
	hipStream_t s = nullptr;
	uint64_t *d_a = nullptr, *d_u = nullptr, *d_ao = nullptr, *d_pk = nullptr;
	int64_t *d_off = nullptr, *d_kept = nullptr, *d_poff = nullptr;
	int32_t *d_nu = nullptr;
	std::vector<int64_t> kept(n_q), poff(2 * n_q);
	std::vector<uint64_t> pk;
	const size_t na = (size_t)(n_a > 0? n_a : 1);
	const nsh_chain::Params prm = { max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, chn_coef_log,
		is_spliced, kmer, bbit };
	std::lock_guard<std::recursive_mutex> lk(g_ws_mu);
	{
		CHOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		CHOK(g_io.need(3 * al256(na * 8)));   // the anchors, the chains, their anchors
		d_a = (uint64_t*)g_io.p, d_u = (uint64_t*)((char*)g_io.p + al256(na * 8));
		d_ao = (uint64_t*)((char*)g_io.p + 2 * al256(na * 8));
		CHOK(hipMalloc((void**)&d_off, (size_t)(n_q + 1) * 8));
		CHOK(hipMalloc((void**)&d_nu, (size_t)n_q * 4));
		if (n_a > 0 && !nsh_copy_to_device(d_a, anchors, (size_t)n_a * 8, s)) goto done;
		CHOK(hipMemcpyAsync(d_off, a_off, (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		if (!nsh_chain::run(s, n_q, d_off, nullptr, cnt.data(), d_a, n_a, prm, d_nu, d_u, d_ao)) goto done;
		// only the chains and their anchors come back, packed, then go to each query's offset
		CHOK(hipMalloc((void**)&d_kept, (size_t)n_q * 8));
		hipLaunchKernelGGL(chain_kept_kernel, dim3((n_q + 255) / 256), dim3(256), 0, s, n_q, d_off, d_nu, d_u, d_kept);
		CHOK(hipGetLastError());
		CHOK(hipMemcpyAsync(n_u, d_nu, (size_t)n_q * 4, hipMemcpyDeviceToHost, s));
		CHOK(hipMemcpyAsync(kept.data(), d_kept, (size_t)n_q * 8, hipMemcpyDeviceToHost, s));
		CHOK(hipStreamSynchronize(s));
		int64_t np = 0;
		for (int32_t q = 0; q < n_q; ++q) poff[q] = np, np += n_u[q];
		for (int32_t q = 0; q < n_q; ++q) poff[n_q + q] = np, np += kept[q];
		if (np > 0) {
			CHOK(hipMalloc((void**)&d_poff, (size_t)n_q * 16));
			CHOK(hipMalloc((void**)&d_pk, (size_t)np * 8));
			CHOK(hipMemcpyAsync(d_poff, poff.data(), (size_t)n_q * 16, hipMemcpyHostToDevice, s));
			hipLaunchKernelGGL(chain_pack_kernel, dim3(n_q), dim3(256), 0, s, d_off, d_nu, d_kept, d_poff,
					d_poff + n_q, d_u, d_ao, d_pk);
			CHOK(hipGetLastError());
			pk.resize(np);
			if (!nsh_copy_to_host(pk.data(), d_pk, (size_t)np * 8, s)) goto done;
			nsh_count_launch(2);
			for (int32_t q = 0; q < n_q; ++q) {
				std::copy(pk.begin() + poff[q], pk.begin() + poff[q] + n_u[q], u + a_off[q]);
				std::copy(pk.begin() + poff[n_q + q], pk.begin() + poff[n_q + q] + kept[q], a_out + a_off[q]);
			}
		}
		total = 0;
		for (int32_t q = 0; q < n_q; ++q) total += n_u[q];
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_off, d_nu, d_kept, d_poff, d_pk };   // d_a, d_u, d_ao: the kept workspace
		for (void *bb : bufs) if (bb) (void)hipFree(bb);
	}
	if (s) (void)hipStreamDestroy(s);
	return total;
}
