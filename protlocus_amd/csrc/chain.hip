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
// Every query's chaining is that sequential procedure, so a lane runs one query from start to
// end (queries ordered by anchor count, so that a wave's lanes finish together), and the batch
// fills the GPU with queries: the same DP, the same backtrack and the same
// radix sort (its recursion unrolled into frames in a per-lane scratch area), in the reference's
// int32 and float32 arithmetic. Per anchor each lane keeps f, p, t and the backtrack list in
// global scratch laid out like the input.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int kRsMin = 64, kRsBits = 8, kRsSize = 1 << kRsBits;
constexpr int kBlockBonus = 2;   // MP_BLOCK_BONUS (protlocus.h:47)

struct ChainParams {
	int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc;
	float coef;
	int32_t is_spliced, kmer, bbit;
};

struct RsFrame {   // one level of the radix sort's recursion
	int32_t s_child, recurse, k;
	int32_t b[kRsSize], e[kRsSize];
};

struct ChArgs {
	const uint64_t *a; const int64_t *off; int32_t n_q;
	const int32_t *order;              // queries by decreasing anchor count (a wave's lanes alike)
	ChainParams cp;
	int32_t *f, *p, *t, *vv;           // per anchor
	uint64_t *zx, *zy;                 // per anchor: sort pairs
	uint64_t *b;                       // per anchor: the chains' anchors before the final order
	RsFrame *frames;                   // per query: 8 frames
	uint64_t *u_out, *a_out;           // per anchor slot: chains (score << 32 | count), their anchors
	int32_t *n_u;                      // per query
};

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

// rs_insertsort (ksort.h:116-126): stable insertion by key
__device__ void rs_insert(uint64_t *x, uint64_t *y, int32_t beg, int32_t end)
{
	for (int32_t i = beg + 1; i < end; ++i)
		if (x[i] < x[i - 1]) {
			const uint64_t tx = x[i], ty = y[i];
			int32_t j = i;
			for (; j > beg && tx < x[j - 1]; --j) x[j] = x[j - 1], y[j] = y[j - 1];
			x[j] = tx, y[j] = ty;
		}
}

// one distribution pass of rs_sort (ksort.h:127-150) over [beg, end) by the digit at bit s
__device__ void rs_distribute(uint64_t *x, uint64_t *y, int32_t beg, int32_t end, int32_t s, RsFrame &F)
{
	const int32_t m = kRsSize - 1;
	for (int32_t k = 0; k < kRsSize; ++k) F.b[k] = F.e[k] = beg;
	for (int32_t i = beg; i < end; ++i) ++F.e[(x[i] >> s) & m];
	for (int32_t k = 1; k < kRsSize; ++k) F.e[k] += F.e[k - 1] - beg, F.b[k] = F.e[k - 1];
	for (int32_t k = 0; k < kRsSize;) {
		if (F.b[k] != F.e[k]) {
			int32_t l = (int32_t)((x[F.b[k]] >> s) & m);
			if (l != k) {
				uint64_t tx = x[F.b[k]], ty = y[F.b[k]];
				do {
					const uint64_t sx = tx, sy = ty;
					tx = x[F.b[l]], ty = y[F.b[l]];
					x[F.b[l]] = sx, y[F.b[l]] = sy;
					++F.b[l];
					l = (int32_t)((tx >> s) & m);
				} while (l != k);
				x[F.b[k]] = tx, y[F.b[k]] = ty;
				++F.b[k];
			} else ++F.b[k];
		} else ++k;
	}
	F.b[0] = beg;
	for (int32_t k = 1; k < kRsSize; ++k) F.b[k] = F.e[k - 1];
	F.recurse = s != 0;
	F.s_child = s > kRsBits? s - kRsBits : 0;
	F.k = 0;
}

// radix_sort_mp128x (ksort.h:158-162) of n (x, y) pairs by x; the recursion runs on frames
__device__ void radix_sort_pairs(uint64_t *x, uint64_t *y, int32_t n, RsFrame *fr)
{
	if (n <= kRsMin) { rs_insert(x, y, 0, n); return; }
	int depth = 0;
	rs_distribute(x, y, 0, n, 7 * kRsBits, fr[0]);
	while (depth >= 0) {
		RsFrame &F = fr[depth];
		if (!F.recurse || F.k >= kRsSize) { --depth; continue; }
		const int32_t k = F.k++, sz = F.e[k] - F.b[k];
		if (sz > kRsMin) {
			rs_distribute(x, y, F.b[k], F.e[k], F.s_child, fr[depth + 1]);
			++depth;
		} else if (sz > 1) rs_insert(x, y, F.b[k], F.e[k]);
	}
}

__global__ void __launch_bounds__(64) chain_kernel(ChArgs A)
{
	const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
	if (tid >= A.n_q) return;
	const int32_t q = A.order[tid];
	const int64_t o = A.off[q], n = A.off[q + 1] - o;
	const uint64_t *a = A.a + o;
	int32_t *f = A.f + o, *p = A.p + o, *t = A.t + o, *vv = A.vv + o;
	uint64_t *zx = A.zx + o, *zy = A.zy + o, *b = A.b + o, *uo = A.u_out + o, *ao = A.a_out + o;
	RsFrame *fr = A.frames + (int64_t)q * 8;
	ChainParams c = A.cp;
	A.n_u[q] = 0;
	if (n == 0) return;
	if (c.max_dist_x < c.bw) c.max_dist_x = c.bw;
	if (c.max_dist_y < c.bw && !c.is_spliced) c.max_dist_y = c.bw;
	const int32_t max_drop = c.is_spliced? INT32_MAX : c.bw;
	// the DP (chain.c:178-208)
	int64_t st = 0, hi = -1;
	int32_t hf = 0;
	for (int64_t i = 0; i < n; ++i) t[i] = 0;
	for (int64_t i = 0; i < n; ++i) {
		int64_t max_j = -1;
		int32_t max_f = c.kmer, n_skip = 0;
		while (st < i && ((a[i] >> 32) - (a[st] >> 32)) << c.bbit > (uint64_t)(int64_t)c.max_dist_x) ++st;
		if (hi >= 0 && hi >= st) {
			const int32_t sc = hf + comput_sc(a[i], a[hi], c);
			if (sc > max_f) max_f = sc, max_j = hi;
		} else hf = 0, hi = -1;
		if (i - st > c.max_iter) st = i - c.max_iter;
		for (int64_t j = i - 1; j >= st; --j) {
			int32_t sc = comput_sc(a[i], a[j], c);
			if (sc == INT32_MIN) continue;
			sc += f[j];
			if (sc > max_f) {
				max_f = sc, max_j = j;
				if (n_skip > 0) --n_skip;
			} else if (t[j] == (int32_t)i) {
				if (++n_skip > c.max_skip) break;
			}
			if (p[j] >= 0) t[p[j]] = (int32_t)i;
		}
		f[i] = max_f, p[i] = (int32_t)max_j;
		if (hf < max_f) hf = max_f, hi = i;
	}
	// the backtrack (mp_chain_backtrack, chain.c:27-82): candidates by score, best first
	int32_t n_z = 0;
	for (int64_t i = 0; i < n; ++i)
		if (f[i] >= c.min_sc) zx[n_z] = (uint64_t)(int64_t)f[i], zy[n_z] = (uint64_t)i, ++n_z;
	if (n_z == 0) return;
	radix_sort_pairs(zx, zy, n_z, fr);
	for (int64_t i = 0; i < n; ++i) t[i] = 0;
	int32_t n_u = 0, n_v = 0;
	for (int32_t k = n_z - 1; k >= 0; --k) {
		const int32_t zk = (int32_t)zy[k], zs = (int32_t)zx[k];
		if (t[zk] != 0) continue;
		// mp_chain_bk_end (chain.c:8-25)
		int32_t end_i = -1, max_i = zk, max_s = 0, ii = zk;
		if (t[ii] == 0) {
			do {
				t[ii] = 2;
				end_i = ii = p[ii];
				const int32_t s = ii < 0? zs : zs - f[ii];
				if (s > max_s) max_s = s, max_i = ii;
				else if (max_s - s > max_drop) break;
			} while (ii >= 0 && t[ii] == 0);
			for (ii = zk; ii >= 0 && ii != end_i; ii = p[ii]) t[ii] = 0;
			end_i = max_i;
		} else end_i = ii;
		const int32_t n_v0 = n_v;
		for (ii = zk; ii != end_i; ii = p[ii]) vv[n_v++] = ii, t[ii] = 1;
		const int32_t sc = ii < 0? zs : zs - f[ii];
		if (sc >= c.min_sc && n_v > n_v0 && n_v - n_v0 >= c.min_cnt)
			uo[n_u++] = (uint64_t)(uint32_t)sc << 32 | (uint32_t)(n_v - n_v0);
		else n_v = n_v0;
	}
	if (n_u == 0) return;
	// compact_a (chain.c:84-110): each chain's anchors in increasing order, then the chains by
	// their first target block
	for (int32_t i = 0, k = 0; i < n_u; ++i) {
		const int32_t k0 = k, ni = (int32_t)(uint32_t)uo[i];
		for (int32_t j = 0; j < ni; ++j) b[k++] = a[vv[k0 + (ni - j - 1)]];
	}
	for (int32_t i = 0, k = 0; i < n_u; ++i) {
		zx[i] = b[k] >> 32, zy[i] = (uint64_t)k << 32 | (uint32_t)i;
		k += (int32_t)(uint32_t)uo[i];
	}
	radix_sort_pairs(zx, zy, n_u, fr);
	// the sorted order: chains into p (as indices), anchors into the output
	for (int32_t i = 0, k = 0; i < n_u; ++i) {
		const int32_t j = (int32_t)(uint32_t)zy[i], cnt = (int32_t)(uint32_t)uo[j];
		const int64_t from = (int64_t)(zy[i] >> 32);
		for (int32_t x = 0; x < cnt; ++x) ao[k + x] = b[from + x];
		k += cnt;
		vv[i] = j;
	}
	for (int32_t i = 0; i < n_u; ++i) zx[i] = uo[vv[i]];
	for (int32_t i = 0; i < n_u; ++i) uo[i] = zx[i];
	A.n_u[q] = n_u;
}

} // namespace

#define CHOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("chain: %s: %s", #x, hipGetErrorString(e_)); goto done; } } while (0)

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
	for (int32_t q = 0; q < n_q; ++q)
		if (a_off[q + 1] < a_off[q] || a_off[q + 1] - a_off[q] > INT32_MAX / 2) {
			nsh_set_error("chain: query %d has a bad anchor range", q);
			return -1;
		}
	if (n_a > 0 && (!anchors || !u || !a_out)) { nsh_set_error("chain: bad arguments"); return -1; }
	if (!nsh_ensure_device()) return -1;
	int64_t total = -1;
	hipStream_t s = nullptr;
	uint64_t *d_a = nullptr, *d_zx = nullptr, *d_zy = nullptr, *d_b = nullptr, *d_u = nullptr, *d_ao = nullptr;
	int64_t *d_off = nullptr;
	int32_t *d_f = nullptr, *d_p = nullptr, *d_t = nullptr, *d_vv = nullptr, *d_nu = nullptr, *d_ord = nullptr;
	std::vector<int32_t> order(n_q);
	RsFrame *d_fr = nullptr;
	const size_t na = (size_t)(n_a > 0? n_a : 1);
	ChArgs A;
	{
		CHOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		CHOK(hipMalloc((void**)&d_a, na * 8));
		CHOK(hipMalloc((void**)&d_off, (size_t)(n_q + 1) * 8));
		CHOK(hipMalloc((void**)&d_f, na * 4));
		CHOK(hipMalloc((void**)&d_p, na * 4));
		CHOK(hipMalloc((void**)&d_t, na * 4));
		CHOK(hipMalloc((void**)&d_vv, na * 4));
		CHOK(hipMalloc((void**)&d_zx, na * 8));
		CHOK(hipMalloc((void**)&d_zy, na * 8));
		CHOK(hipMalloc((void**)&d_b, na * 8));
		CHOK(hipMalloc((void**)&d_u, na * 8));
		CHOK(hipMalloc((void**)&d_ao, na * 8));
		CHOK(hipMalloc((void**)&d_nu, (size_t)n_q * 4));
		CHOK(hipMalloc((void**)&d_fr, (size_t)n_q * 8 * sizeof(RsFrame)));
		CHOK(hipMalloc((void**)&d_ord, (size_t)n_q * 4));
		for (int32_t q = 0; q < n_q; ++q) order[q] = q;
		std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
			return a_off[x + 1] - a_off[x] > a_off[y + 1] - a_off[y]; });
		CHOK(hipMemcpyAsync(d_ord, order.data(), (size_t)n_q * 4, hipMemcpyHostToDevice, s));
		if (n_a > 0) CHOK(hipMemcpyAsync(d_a, anchors, (size_t)n_a * 8, hipMemcpyHostToDevice, s));
		CHOK(hipMemcpyAsync(d_off, a_off, (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		A.a = d_a, A.off = d_off, A.n_q = n_q, A.order = d_ord;
		A.cp = { max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, chn_coef_log, is_spliced, kmer, bbit };
		A.f = d_f, A.p = d_p, A.t = d_t, A.vv = d_vv, A.zx = d_zx, A.zy = d_zy, A.b = d_b, A.frames = d_fr;
		A.u_out = d_u, A.a_out = d_ao, A.n_u = d_nu;
		hipLaunchKernelGGL(chain_kernel, dim3((n_q + 63) / 64), dim3(64), 0, s, A);
		CHOK(hipGetLastError());
		CHOK(hipMemcpyAsync(n_u, d_nu, (size_t)n_q * 4, hipMemcpyDeviceToHost, s));
		if (n_a > 0) {
			CHOK(hipMemcpyAsync(u, d_u, (size_t)n_a * 8, hipMemcpyDeviceToHost, s));
			CHOK(hipMemcpyAsync(a_out, d_ao, (size_t)n_a * 8, hipMemcpyDeviceToHost, s));
		}
		CHOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		total = 0;
		for (int32_t q = 0; q < n_q; ++q) total += n_u[q];
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_a, d_off, d_f, d_p, d_t, d_vv, d_zx, d_zy, d_b, d_u, d_ao, d_nu, d_fr, d_ord };
		for (void *bb : bufs) if (bb) (void)hipFree(bb);
	}
	if (s) (void)hipStreamDestroy(s);
	return total;
}
