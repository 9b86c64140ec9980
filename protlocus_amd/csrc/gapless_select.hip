// The --gpu-prefilter candidate selection for a whole mini-batch on the device: the batch
// calibration (mp_gapless_calibrate, hit.c:149-189), the per-query K (mp_gapless_adaptive_k,
// 195-215) and the top-K selection (mp_select_gapless, 217-263, with its small-limit path
// 87-136), driven as mp_gpu_prefilter_batch / mp_gpu_prefilter_select drive them
// (map.c:1075-1083, 507-521).
//
// The reference ranks each query's candidates with qsort (or a streaming top-8), then walks the
// ranking. Here one wave owns a query:
//   * every candidate's rank in the total order (gapless score desc, chain score desc, vid asc,
//     vs asc, index asc) and its rank among candidates of its own vid are counted against all of
//     the query's candidates;
//   * a candidate is eligible for the per-vid floor when its vid rank is below min_per_vid; the
//     walk keeps the first `limit` eligible candidates in rank order, then fills with the best
//     ineligible ones. Both counts come from one more pass over the ranks held in LDS. This is
//     what both of the reference's paths select: its streaming top-8 keeps, after each
//     candidate, the greedy selection of the candidates seen so far;
//   * the chain guard (hit.c:71-85) comes from two wave arg-max reductions.
// The calibration's background median is a device select + radix sort; the duplication counts
// are wave reductions and three atomics.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int kTiePct = 80;          // MP_GAPLESS_TIE_PCT (mppriv.h:94)
constexpr float kDupRich = 0.25f;    // MP_GAPLESS_DUP_RICH (mppriv.h:95)
constexpr int kAdaptMaxK = 4;        // MP_GAPLESS_ADAPT_MAX_K (mppriv.h:96)
constexpr int kMaxPerQuery = 8192;   // candidates of one query (LDS: 4 bytes each)

struct SelArgs {
	const int32_t *off, *qlen, *sc, *chn;
	const uint32_t *vid;
	const int64_t *vs;
	int32_t n_query, top_k, min_per_vid, adaptive, noise, dup_rich;
	uint8_t *keep, *bg;
	int32_t *kq, *cnt;                // cnt: queries scored, confident, with a near tie
};

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d));
	return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
	return v;
}

// the query's best score and its runner-up, a tied winner counting as its own runner-up
// (hit.c:157-163)
__device__ void best2(const int32_t *sc, int n, int L, int &best, int &second)
{
	int b = INT_MIN;
	for (int i = L; i < n; i += 64) b = max(b, sc[i]);
	best = wave_max(b);
	int nb = 0, s = INT_MIN;
	for (int i = L; i < n; i += 64) {
		nb += sc[i] == best;
		if (sc[i] < best) s = max(s, sc[i]);
	}
	nb = wave_sum(nb);
	second = nb >= 2? best : wave_max(s);
}

// calibration, first pass: the background flags (0 < score < the query's best) and the count of
// queries that contribute
__global__ void __launch_bounds__(64) calib_bg_kernel(SelArgs a)
{
	const int q = blockIdx.x, L = threadIdx.x;
	const int o = a.off[q], n = a.off[q + 1] - o;
	const int32_t *sc = a.sc + o;
	if (n < 2) {
		for (int i = L; i < n; i += 64) a.bg[o + i] = 0;
		return;
	}
	int best, second;
	best2(sc, n, L, best, second);
	for (int i = L; i < n; i += 64) a.bg[o + i] = best > 0 && sc[i] > 0 && sc[i] < best;
	if (L == 0 && best > 0) atomicAdd(&a.cnt[0], 1);
}

// calibration, second pass: queries whose winner clears 4 x noise, and those with a near tie
__global__ void __launch_bounds__(64) calib_dup_kernel(SelArgs a)
{
	const int q = blockIdx.x, L = threadIdx.x;
	const int o = a.off[q], n = a.off[q + 1] - o;
	if (n < 2) return;
	int best, second;
	best2(a.sc + o, n, L, best, second);
	if (best <= 0 || best < 4 * a.noise) return;
	if (L == 0) {
		atomicAdd(&a.cnt[1], 1);
		if (second > 0 && (int64_t)second * 100 >= (int64_t)best * kTiePct) atomicAdd(&a.cnt[2], 1);
	}
}

// candidate j ranks before candidate i (mp_gapless_rank_cmp, hit.c:42-53)
__device__ __forceinline__ bool before(int sj, int cj, uint32_t vj, int64_t xj, int j, int si, int ci, uint32_t vi,
		int64_t xi, int i)
{
	if (sj != si) return sj > si;
	if (cj != ci) return cj > ci;
	if (vj != vi) return vj < vi;
	if (xj != xi) return xj < xi;
	return j < i;
}

// one wave per query: K, then the kept candidates
__global__ void __launch_bounds__(64) select_kernel(SelArgs a)
{
	extern __shared__ int32_t rk[];     // per candidate: rank << 1 | eligible
	const int q = blockIdx.x, L = threadIdx.x;
	const int o = a.off[q], n = a.off[q + 1] - o;
	if (n <= 0) {
		if (L == 0) a.kq[q] = 0;
		return;
	}
	const int32_t *sc = a.sc + o, *chn = a.chn + o;
	const uint32_t *vid = a.vid + o;
	const int64_t *vs = a.vs + o;
	int k = a.top_k;
	if (a.adaptive) {   // mp_gapless_adaptive_k (hit.c:195-215)
		const int max_k = min(k, kAdaptMaxK);
		int b = INT_MIN;
		for (int i = L; i < n; i += 64) b = max(b, sc[i]);
		const int best = wave_max(b);
		int target;
		if (best <= 0 || best < 2 * a.noise || (int64_t)best * 4 < a.qlen[q]) {
			target = max_k;
		} else {
			int64_t thr = (int64_t)best * kTiePct / 100;
			if (thr < 2 * (int64_t)a.noise) thr = 2 * (int64_t)a.noise;
			int c = 0;
			for (int i = L; i < n; i += 64) c += sc[i] >= thr;
			target = wave_sum(c) + a.dup_rich;
		}
		k = max(1, min(min(target, max_k), n));
	}
	if (L == 0) a.kq[q] = k;
	const int limit = min(n, k);
	for (int i = L; i < n; i += 64) {
		const int si = sc[i], ci = chn[i];
		const uint32_t vi = vid[i];
		const int64_t xi = vs[i];
		int r = 0, vr = 0;
		for (int j = 0; j < n; ++j) {
			const bool bj = before(sc[j], chn[j], vid[j], vs[j], j, si, ci, vi, xi, i);
			r += bj, vr += bj && vid[j] == vi;
		}
		rk[i] = r << 1 | (a.min_per_vid > 0 && vr < a.min_per_vid);
	}
	__syncthreads();
	int ne = 0;
	for (int i = L; i < n; i += 64) ne += rk[i] & 1;
	ne = wave_sum(ne);
	// the chain guard: the best chain (chain score, then gapless score, first index) against the
	// gapless winner (score, then chain score, first index), when the selection drops candidates
	int guard = -1;
	if (limit < n) {
		int bi = -1, ti = -1;
		for (int i = L; i < n; i += 64) {
			if (bi < 0 || chn[i] > chn[bi] || (chn[i] == chn[bi] && sc[i] > sc[bi])) bi = i;
			if (ti < 0 || sc[i] > sc[ti] || (sc[i] == sc[ti] && chn[i] > chn[ti])) ti = i;
		}
#pragma unroll
		for (int d = 32; d >= 1; d >>= 1) {
			const int ob = __shfl_xor(bi, d), ot = __shfl_xor(ti, d);
			if (ob >= 0 && (bi < 0 || chn[ob] > chn[bi] || (chn[ob] == chn[bi] && (sc[ob] > sc[bi] ||
					(sc[ob] == sc[bi] && ob < bi))))) bi = ob;
			if (ot >= 0 && (ti < 0 || sc[ot] > sc[ti] || (sc[ot] == sc[ti] && (chn[ot] > chn[ti] ||
					(chn[ot] == chn[ti] && ot < ti))))) ti = ot;
		}
		if (bi != ti && (int64_t)chn[ti] * 10 < (int64_t)chn[bi] * 8) guard = bi;
	}
	for (int i = L; i < n; i += 64) {
		const int ri = rk[i] >> 1, ei = rk[i] & 1;
		int e = 0, x = 0;   // eligible / ineligible candidates ranked before i
		for (int j = 0; j < n; ++j) {
			const int v = rk[j];
			if ((v >> 1) < ri) {
				if (v & 1) ++e;
				else ++x;
			}
		}
		const bool sel = ei? e < limit : (ne < limit && x < limit - ne);
		a.keep[o + i] = sel || i == guard;
	}
}

__global__ void zero_kernel(int32_t *p, int n)
{
	if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// the entry points share the workspace: nsh_gapless_select_batch holds the lock across its call
// of nsh_gsel::run, which takes it again
std::recursive_mutex g_mu;
hipStream_t g_stream = nullptr;
struct Buf { void *p = nullptr; size_t n = 0; };
enum { B_OFF, B_QLEN, B_SC, B_CHN, B_VID, B_VS, B_KEEP, B_K, B_BG, B_BGV, B_BGS, B_CNT, B_TMP, B_PIN, B_N };
Buf g_buf[B_N];

// grows buffer b; the work queued on stream s (the only user of the workspace while the lock is
// held) finishes first
hipError_t grow(int b, size_t bytes, hipStream_t s, bool pinned = false)
{
	bytes += 64;
	if (g_buf[b].n >= bytes) return hipSuccess;
	{
		const hipError_t e = hipStreamSynchronize(s);
		if (e != hipSuccess) return e;
	}
	if (g_buf[b].p) (void)(pinned? hipHostFree(g_buf[b].p) : hipFree(g_buf[b].p));
	g_buf[b].p = nullptr, g_buf[b].n = 0;
	const hipError_t e = pinned? hipHostMalloc(&g_buf[b].p, bytes, hipHostMallocDefault) : hipMalloc(&g_buf[b].p, bytes);
	if (e == hipSuccess) g_buf[b].n = bytes;
	return e;
}

#define GSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("gapless select: %s: %s", #x, hipGetErrorString(e_)); return 0; } } while (0)

} // namespace

// device-resident inputs on stream s (the fused scoring + selection, and the host entry below)
int nsh_gsel::run(hipStream_t s, int32_t n_query, const int32_t *d_off, const int32_t *d_qlen, const int32_t *d_sc,
		const int32_t *d_chn, const uint32_t *d_vid, const int64_t *d_vs, int32_t n_cand, int32_t max_per_query,
		int32_t top_k, int32_t min_per_vid, int32_t adaptive, uint8_t *d_keep, int32_t *d_k, nsh_gapless_calib_t *cal)
{
	std::lock_guard<std::recursive_mutex> lk(g_mu);
	SelArgs a;
	a.off = d_off, a.qlen = d_qlen, a.sc = d_sc, a.chn = d_chn, a.vid = d_vid, a.vs = d_vs;
	a.n_query = n_query, a.top_k = top_k, a.min_per_vid = min_per_vid, a.adaptive = adaptive;
	a.noise = 0, a.dup_rich = 0, a.keep = d_keep, a.kq = d_k, a.bg = nullptr, a.cnt = nullptr;
	nsh_gapless_calib_t c = { 0, 0, 0.0f };
	if (n_query == 0) { if (cal) *cal = c; return 1; }
	if (adaptive) {   // mp_gapless_calibrate (hit.c:149-189)
		GSOK(grow(B_BG, (size_t)n_cand, s));
		GSOK(grow(B_BGV, (size_t)n_cand * 4, s));
		GSOK(grow(B_BGS, (size_t)n_cand * 4, s));
		GSOK(grow(B_CNT, 16, s));
		GSOK(grow(B_PIN, 32, s, true));
		a.bg = (uint8_t*)g_buf[B_BG].p, a.cnt = (int32_t*)g_buf[B_CNT].p;
		int32_t *bgv = (int32_t*)g_buf[B_BGV].p, *bgs = (int32_t*)g_buf[B_BGS].p, *cnt = a.cnt;
		int32_t *pin = (int32_t*)g_buf[B_PIN].p;
		// counters: scored, confident, near tie, and (slot 3) the number of background scores
		hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, cnt, 4);
		hipLaunchKernelGGL(calib_bg_kernel, dim3(n_query), dim3(64), 0, s, a);
		size_t need = 0, t2 = 0;
		GSOK(hipcub::DeviceSelect::Flagged(nullptr, need, d_sc, a.bg, bgv, cnt + 3, n_cand, s));
		GSOK(hipcub::DeviceRadixSort::SortKeys(nullptr, t2, bgv, bgs, n_cand, 0, 32, s));
		GSOK(grow(B_TMP, std::max(need, t2), s));
		need = g_buf[B_TMP].n;
		GSOK(hipcub::DeviceSelect::Flagged(g_buf[B_TMP].p, need, d_sc, a.bg, bgv, cnt + 3, n_cand, s));
		GSOK(hipMemcpyAsync(pin, cnt, 16, hipMemcpyDeviceToHost, s));
		GSOK(hipStreamSynchronize(s));
		const int32_t n_bg = pin[3];
		c.n_query = pin[0];
		if (n_bg >= 64) {
			need = g_buf[B_TMP].n;
			GSOK(hipcub::DeviceRadixSort::SortKeys(g_buf[B_TMP].p, need, bgv, bgs, n_bg, 0, 32, s));
			GSOK(hipMemcpyAsync(pin + 4, bgs + n_bg / 2, 4, hipMemcpyDeviceToHost, s));
			GSOK(hipStreamSynchronize(s));
			c.noise = pin[4];
		}
		a.noise = c.noise;
		hipLaunchKernelGGL(calib_dup_kernel, dim3(n_query), dim3(64), 0, s, a);
		GSOK(hipMemcpyAsync(pin, cnt, 16, hipMemcpyDeviceToHost, s));
		GSOK(hipStreamSynchronize(s));
		c.dup_frac = pin[1] > 0? (float)pin[2] / pin[1] : 0.0f;
		a.dup_rich = c.dup_frac >= kDupRich;
		nsh_count_launch(4);
	}
	hipLaunchKernelGGL(select_kernel, dim3(n_query), dim3(64), (size_t)max_per_query * 4, s, a);
	GSOK(hipGetLastError());
	nsh_count_launch(1);
	if (cal) *cal = c;
	return 1;
}

extern "C" int nsh_gapless_select_batch(int32_t n_query, const int32_t *cand_off, const int32_t *qlen,
		const int32_t *scores, const int32_t *chn_sc, const uint32_t *vid, const int64_t *vs, int32_t top_k,
		int32_t min_per_vid, int32_t adaptive, uint8_t *keep, int32_t *k_out, nsh_gapless_calib_t *cal)
{
	if (n_query < 0 || (n_query > 0 && (!cand_off || !qlen)) || top_k < 1 || min_per_vid < 0 || min_per_vid > top_k) {
		nsh_set_error("gapless select: bad arguments (need top_k >= 1, 0 <= min_per_vid <= top_k)");
		return 0;
	}
	int32_t maxn = 0;
	if (n_query > 0 && cand_off[0] != 0) { nsh_set_error("gapless select: cand_off[0] must be 0"); return 0; }
	for (int32_t q = 0; q < n_query; ++q) {
		const int32_t n = cand_off[q + 1] - cand_off[q];
		if (n < 0 || n > kMaxPerQuery) {
			nsh_set_error("gapless select: query %d has %d candidates (0..%d)", q, n, kMaxPerQuery);
			return 0;
		}
		maxn = std::max(maxn, n);
	}
	const int32_t n_cand = n_query > 0? cand_off[n_query] : 0;
	if (n_cand > 0 && (!scores || !chn_sc || !vid || !vs || !keep)) { nsh_set_error("gapless select: bad arguments"); return 0; }
	if (!nsh_ensure_device()) return 0;
	std::lock_guard<std::recursive_mutex> lk(g_mu);
	if (!g_stream) GSOK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
	hipStream_t s = g_stream;
	GSOK(grow(B_OFF, (size_t)(n_query + 1) * 4, s));
	GSOK(grow(B_QLEN, (size_t)n_query * 4, s));
	GSOK(grow(B_SC, (size_t)n_cand * 4, s));
	GSOK(grow(B_CHN, (size_t)n_cand * 4, s));
	GSOK(grow(B_VID, (size_t)n_cand * 4, s));
	GSOK(grow(B_VS, (size_t)n_cand * 8, s));
	GSOK(grow(B_KEEP, (size_t)n_cand, s));
	GSOK(grow(B_K, (size_t)n_query * 4, s));
	if (n_query > 0) {
		GSOK(hipMemcpyAsync(g_buf[B_OFF].p, cand_off, (size_t)(n_query + 1) * 4, hipMemcpyHostToDevice, s));
		GSOK(hipMemcpyAsync(g_buf[B_QLEN].p, qlen, (size_t)n_query * 4, hipMemcpyHostToDevice, s));
	}
	if (n_cand > 0) {
		GSOK(hipMemcpyAsync(g_buf[B_SC].p, scores, (size_t)n_cand * 4, hipMemcpyHostToDevice, s));
		GSOK(hipMemcpyAsync(g_buf[B_CHN].p, chn_sc, (size_t)n_cand * 4, hipMemcpyHostToDevice, s));
		GSOK(hipMemcpyAsync(g_buf[B_VID].p, vid, (size_t)n_cand * 4, hipMemcpyHostToDevice, s));
		GSOK(hipMemcpyAsync(g_buf[B_VS].p, vs, (size_t)n_cand * 8, hipMemcpyHostToDevice, s));
	}
	if (!nsh_gsel::run(s, n_query, (const int32_t*)g_buf[B_OFF].p, (const int32_t*)g_buf[B_QLEN].p,
			(const int32_t*)g_buf[B_SC].p, (const int32_t*)g_buf[B_CHN].p, (const uint32_t*)g_buf[B_VID].p,
			(const int64_t*)g_buf[B_VS].p, n_cand, maxn, top_k, min_per_vid, adaptive, (uint8_t*)g_buf[B_KEEP].p,
			(int32_t*)g_buf[B_K].p, cal))
		return 0;
	if (n_cand > 0) GSOK(hipMemcpyAsync(keep, g_buf[B_KEEP].p, (size_t)n_cand, hipMemcpyDeviceToHost, s));
	if (k_out && n_query > 0) GSOK(hipMemcpyAsync(k_out, g_buf[B_K].p, (size_t)n_query * 4, hipMemcpyDeviceToHost, s));
	GSOK(hipStreamSynchronize(s));
	return 1;
}
