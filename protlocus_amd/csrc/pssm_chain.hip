// PSSM diagonal evidence of refine chains: mp_diag_seg and mp_chain_pssm_ev (map.c:91-139).
//
// One wave per chain. The chain's anchors are taken in order (the duplicate test depends on
// the segments recorded before), so the anchor loop is wave-uniform. Each anchor's diagonal walk
// is the data-parallel part: 64 positions per step, one per lane. A lane scores its codon against
// the PSSM row, then a wave scan builds the running sum and the running best. The walk stops at
// the first position that is not a new best and lies more than xdrop below the best, exactly as
// the sequential loop does. Up to 64 recorded segments (diagonal, protein end) live one per lane,
// and the duplicate test is a ballot.
#include <hip/hip_runtime.h>
#include <chrono>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

struct ChainArgs {
	const uint8_t *nt; const int64_t *nt_off, *l_nt;       // region bases (codes 0..4) per chain
	const int8_t *pssm; const int64_t *pssm_off; const int32_t *qlen;
	const uint64_t *a; const int64_t *a_off; const int32_t *n_a;
	const uint8_t *codon; int32_t aa_x;                     // codon table; residue of codons with N
	int32_t n, kmer, min_seg, xdrop;
	int32_t *ev, *seg_sc, *seg_e;
};

#define DEV static __device__ __forceinline__

DEV int scan_add(int v, int L)
{
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const int o = __shfl_up(v, d);
		if (L >= d) v += o;
	}
	return v;
}

DEV int scan_max(int v, int L)
{
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const int o = __shfl_up(v, d);
		if (L >= d) v = max(v, o);
	}
	return v;
}

// Positions j0, j0 + dir, ... (npos of them); the base of position j is p0 + 3 (j - q0). Returns
// the best running sum (>= 0) and, in *best_j, the position where it was first reached
// (INT_MIN when it never rose above 0): the loops of mp_diag_seg (map.c:97-120).
DEV int walk(const ChainArgs &a, const uint8_t *nt, const int8_t *pssm, int64_t p0, int q0, int j0, int dir, int npos,
		int *best_j, int L)
{
	int sc = 0, best = 0, bj = INT_MIN;
	for (int base = 0; base < npos; base += 64) {
		const int t = base + L;
		const bool valid = t < npos;
		const int j = j0 + dir * t;
		int s = 0;
		if (valid) {
			const int64_t p = p0 + 3 * (int64_t)(j - q0);
			const int b0 = nt[p], b1 = nt[p + 1], b2 = nt[p + 2];
			const int aa = b0 > 3 || b1 > 3 || b2 > 3? a.aa_x : a.codon[b0 << 4 | b1 << 2 | b2];
			s = pssm[(int64_t)j * 22 + aa];
		}
		const int P = scan_add(s, L) + sc;                     // running sum after position t
		const int M = max(scan_max(valid? P : INT_MIN, L), best);   // running best after t
		int Mex = __shfl_up(M, 1);                             // running best before t
		if (L == 0) Mex = best;
		const uint64_t stop = __ballot(valid && P <= Mex && Mex - P > a.xdrop);
		const int last = stop? __ffsll((unsigned long long)stop) - 1 : min(63, npos - base - 1);
		const int Mlast = __shfl(M, last);
		const uint64_t up = __ballot(valid && L <= last && P > Mex && P == Mlast);   // the last strict new best
		if (up) bj = j0 + dir * (base + __ffsll((unsigned long long)up) - 1);
		best = Mlast, sc = __shfl(P, last);
		if (stop) break;
	}
	*best_j = bj;
	return best;
}

__global__ void __launch_bounds__(256) chain_ev_kernel(ChainArgs a)
{
	const int L = threadIdx.x & 63;
	const int c = blockIdx.x * 4 + (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	if (c >= a.n) return;   // wave-uniform
	const uint8_t *nt = a.nt + a.nt_off[c];
	const int64_t l_nt = a.l_nt[c];
	const int8_t *pssm = a.pssm + a.pssm_off[c];
	const int qlen = a.qlen[c], kmer = a.kmer, na = a.n_a[c];
	const uint64_t *A = a.a + a.a_off[c];
	int m = 0, ev = 0, se = 0;
	int64_t sd = 0;   // lane t < m: recorded segment t (diagonal, protein end)
	for (int i = 0; i < na; ++i) {
		const uint64_t av = A[i];
		const int64_t x = (int64_t)(av >> 32);
		const int q = (int32_t)(uint32_t)av, q0 = q + 1 - kmer;
		const int64_t p0 = x + 1 - 3 * (int64_t)kmer, d = p0 - 3 * (int64_t)q0;
		const bool dup = __ballot(L < m && sd == d && q0 <= se) != 0;
		int sc = INT_MIN, e = -1;
		if (!dup) {
			// forward from q0: positions with a base >= 0 and a whole codon inside the region
			const int64_t D = l_nt - 2 - p0;
			const int64_t fs = p0 >= 0? q0 : q0 + (-p0 + 2) / 3;
			const int64_t fe = D > 0? min((int64_t)qlen, (int64_t)q0 + (D + 2) / 3) : (int64_t)q0;
			int bj;
			const int tot = walk(a, nt, pssm, p0, q0, (int)fs, 1, fe > fs? (int)(fe - fs) : 0, &bj, L);
			e = (bj == INT_MIN? q0 - 1 : bj) + 1;
			// backward from q0 - 1: skip codons past the region end, stop at the region start
			int back = 0;
			if (p0 >= 0 && q0 > 0) {
				const int64_t E = p0 + 2 - l_nt;
				const int64_t top = E < 0? (int64_t)q0 - 1 : (int64_t)q0 - (E / 3 + 1);
				const int64_t bot = max((int64_t)0, (int64_t)q0 - p0 / 3);
				if (top >= bot) back = walk(a, nt, pssm, p0, q0, (int)top, -1, (int)(top - bot + 1), &bj, L);
			}
			sc = tot + back;
			if (m < 64) {
				if (L == m) sd = d, se = e;
				++m;
			}
			if (sc >= a.min_seg) ev += sc;
		}
		if (a.seg_sc && L == 0) a.seg_sc[a.a_off[c] + i] = sc, a.seg_e[a.a_off[c] + i] = e;
	}
	if (L == 0) a.ev[c] = ev;
}

std::mutex g_mu;

} // namespace

#define PCOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("chain_pssm_ev: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" int nsh_chain_pssm_ev_batch(int32_t n, const uint8_t *nt, int64_t l_nt_all, const int64_t *nt_off,
		const int64_t *l_nt, const int8_t *pssm, int64_t l_pssm, const int64_t *pssm_off, const int32_t *qlen,
		const uint64_t *a, int64_t n_a_all, const int64_t *a_off, const int32_t *n_a, int32_t kmer, int32_t min_seg,
		int32_t xdrop, int32_t *ev, int32_t *seg_sc, int32_t *seg_e)
{
	if (!nsh_ensure_device()) return 0;
	if (n < 0 || kmer < 1 || (n > 0 && (!nt || !nt_off || !l_nt || !pssm || !pssm_off || !qlen || !a_off || !n_a || !ev ||
			(n_a_all > 0 && !a))) || (!seg_sc) != (!seg_e)) {
		nsh_set_error("chain_pssm_ev: bad arguments");
		return 0;
	}
	if (n == 0) return 1;
	for (int32_t c = 0; c < n; ++c) {
		if (nt_off[c] < 0 || l_nt[c] < 0 || nt_off[c] + l_nt[c] > l_nt_all || qlen[c] < 0 || pssm_off[c] < 0 ||
				pssm_off[c] + (int64_t)qlen[c] * 22 > l_pssm || a_off[c] < 0 || n_a[c] < 0 || a_off[c] + n_a[c] > n_a_all) {
			nsh_set_error("chain_pssm_ev: chain %d out of range", c);
			return 0;
		}
		for (int32_t i = 0; i < n_a[c]; ++i) {
			const int32_t q = (int32_t)(uint32_t)a[a_off[c] + i];
			if (q + 1 - kmer < 0 || q + 1 - kmer > qlen[c]) {
				nsh_set_error("chain_pssm_ev: chain %d anchor %d: protein end %d outside [kmer - 1, qlen]", c, i, q);
				return 0;
			}
		}
	}
	std::lock_guard<std::mutex> lk(g_mu);
	static hipStream_t stream = nullptr;
	uint8_t *d_nt = nullptr, *d_codon = nullptr;
	int8_t *d_pssm = nullptr;
	int64_t *d_i64 = nullptr;       // nt_off, l_nt, pssm_off, a_off
	int32_t *d_i32 = nullptr;       // qlen, n_a, ev
	uint64_t *d_a = nullptr;
	int32_t *d_seg = nullptr;
	int ok = 0;
	const auto t0 = std::chrono::steady_clock::now();
	{
		if (!stream) PCOK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		PCOK(hipMalloc((void**)&d_nt, (size_t)l_nt_all + 4));
		PCOK(hipMalloc((void**)&d_pssm, (size_t)l_pssm + 4));
		PCOK(hipMalloc((void**)&d_codon, 64));
		PCOK(hipMalloc((void**)&d_i64, (size_t)n * 4 * 8));
		PCOK(hipMalloc((void**)&d_i32, (size_t)n * 3 * 4));
		PCOK(hipMalloc((void**)&d_a, (size_t)n_a_all * 8 + 8));
		if (seg_sc) PCOK(hipMalloc((void**)&d_seg, (size_t)n_a_all * 8 + 8));
		if (l_nt_all) PCOK(hipMemcpyAsync(d_nt, nt, (size_t)l_nt_all, hipMemcpyHostToDevice, stream));
		if (l_pssm) PCOK(hipMemcpyAsync(d_pssm, pssm, (size_t)l_pssm, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_codon, ns_tab_codon, 64, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i64, nt_off, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i64 + n, l_nt, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i64 + 2 * (size_t)n, pssm_off, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i64 + 3 * (size_t)n, a_off, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i32, qlen, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		PCOK(hipMemcpyAsync(d_i32 + n, n_a, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		if (n_a_all) PCOK(hipMemcpyAsync(d_a, a, (size_t)n_a_all * 8, hipMemcpyHostToDevice, stream));
		ChainArgs ca;
		ca.nt = d_nt, ca.nt_off = d_i64, ca.l_nt = d_i64 + n, ca.pssm = d_pssm, ca.pssm_off = d_i64 + 2 * (size_t)n;
		ca.qlen = d_i32, ca.a = d_a, ca.a_off = d_i64 + 3 * (size_t)n, ca.n_a = d_i32 + n;
		ca.codon = d_codon, ca.aa_x = ns_tab_aa20[(uint8_t)'X'];
		ca.n = n, ca.kmer = kmer, ca.min_seg = min_seg, ca.xdrop = xdrop;
		ca.ev = d_i32 + 2 * (size_t)n, ca.seg_sc = d_seg, ca.seg_e = d_seg? d_seg + n_a_all : nullptr;
		hipLaunchKernelGGL(chain_ev_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, ca);
		PCOK(hipGetLastError());
		PCOK(hipMemcpyAsync(ev, d_i32 + 2 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
		if (seg_sc && n_a_all) {
			PCOK(hipMemcpyAsync(seg_sc, d_seg, (size_t)n_a_all * 4, hipMemcpyDeviceToHost, stream));
			PCOK(hipMemcpyAsync(seg_e, d_seg + n_a_all, (size_t)n_a_all * 4, hipMemcpyDeviceToHost, stream));
		}
		PCOK(hipStreamSynchronize(stream));
		nsh_count_launch(1);
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_nt, d_codon, d_pssm, d_i64, d_i32, d_a, d_seg };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}
