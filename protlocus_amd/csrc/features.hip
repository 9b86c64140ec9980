// Chain-candidate features for the learned prescreen (no reference counterpart).
//
// One work-group of 512 threads per candidate (protein aa codes, DNA window as base codes).
// Each string is read from HBM once, coalesced: the window becomes one residue code per base
// position (the codon starting there) and the protein its residue codes, both staged in LDS
// with the score matrix, so every later pass (anchors, composition, segment scores) runs from
// LDS. Anchors are exact amino-acid 3-mer matches between the protein and the three forward
// translation frames of the window; the protein's 3-mers are indexed in LDS by code
// (20^3 = 8000 entries, first position), so each window codon is one LDS lookup. Scalar counts
// are kept per thread and summed per wave; histograms use LDS atomics; one thread turns the
// counts into the 64 features (layout below). Every count is an integer, every feature one
// float expression of them, so a CPU restatement reproduces the features exactly (up to the
// final float rounding).
//
// Features (float32): 0 log2(1+al) | 1 log2(1+nl) | 2 nl/(3al) | 3-5 anchors per frame / al |
// 6 anchors / al | 7-9 stop codons per frame / codons of the frame | 10 codons with N / codons |
// 11 GC / nl | 12 GT / nl * 16 | 13 AG / nl * 16 | 14 anchors of the dominant frame / anchors |
// 15 largest diagonal bin (8 codons wide, per frame) / al | 16 bins with >= 3 anchors / 10 |
// 17 protein positions hit / al | 18 codon positions hit / (nl/3) | 19 sd(diagonal) / al |
// 20-39 protein composition | 40-43 best ungapped segment scores (score matrix) of the four
// fullest diagonal bins, descending, / al | 44 their sum / al | 45 how many reach 20, / 4 |
// 46-59 dominant-frame composition of residues 0-13 | 60-63 anchors per query quarter / al
#include <hip/hip_runtime.h>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include "nsh_internal.h"
#include "prep.h"

namespace {

constexpr int NF = 64;
constexpr int NT = 512;               // threads per candidate (eight waves)
constexpr int KMAX = 8000;            // 20^3
constexpr int AL_MAX = 4096, NL_MAX = 24000;
constexpr int NBIN = (AL_MAX + NL_MAX / 3) / 8 + 2;   // diagonal bins per frame

struct FeatArgs {
	const nsh_prep::PrepProb *pp;   // per candidate: ns_off, as_off, nl, al (raw strings)
	const uint8_t *ns, *as;         // raw DNA (ASCII or 0..4) and protein (ASCII) bytes
	const nsh_prep::PrepConst *pc;  // nt4 / aa20 / codon tables
	const int8_t *mat;              // 22 x 22 scores, [target residue][protein residue]
	int32_t n;
	float *out;                     // n x 64
};

// wave-wide reductions with DPP: quad swaps, half-row and row mirrors leave every lane of a
// 16-lane row with the row's result; the four rows are combined from lanes 0, 16, 32, 48.
// Every lane of the wave must be active.
template <int CTRL> static __device__ __forceinline__ int dpp(int v)
{
	return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
static __device__ __forceinline__ int wave_sum(int v)
{
	v += dpp<0xb1>(v), v += dpp<0x4e>(v), v += dpp<0x141>(v), v += dpp<0x140>(v);
	return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
		__builtin_amdgcn_readlane(v, 48);
}
static __device__ __forceinline__ unsigned wave_umax(unsigned v)
{
	v = max(v, (unsigned)dpp<0xb1>((int)v)), v = max(v, (unsigned)dpp<0x4e>((int)v));
	v = max(v, (unsigned)dpp<0x141>((int)v)), v = max(v, (unsigned)dpp<0x140>((int)v));
	const unsigned a = __builtin_amdgcn_readlane((int)v, 0), b = __builtin_amdgcn_readlane((int)v, 16);
	const unsigned c = __builtin_amdgcn_readlane((int)v, 32), d = __builtin_amdgcn_readlane((int)v, 48);
	return max(max(a, b), max(c, d));
}

__global__ void __launch_bounds__(NT) feature_kernel(FeatArgs a)
{
	__shared__ int first[KMAX];
	__shared__ unsigned hist[3][(NBIN + 1) / 2];   // two 16-bit bins per word
	__shared__ unsigned qhit[AL_MAX / 32], thit[NL_MAX / 3 / 32 + 1];
	__shared__ int cnt[96];   // 0-2 anchors/frame, 3-5 stops/frame, 6 X codons, 7 GC, 8 GT, 9 AG, 10-13 quarters,
	                          // 14-33 protein comp, 34-53 dominant comp, 54 bins>=3, 55 max bin
	__shared__ long long dsum[2];
	__shared__ uint8_t nt4[256], aa20[256], codon[64];
	__shared__ uint8_t cod[NL_MAX];     // cod[i]: residue of the codon at bases i..i+2 (21 if any is not ACGT)
	__shared__ uint8_t pa[AL_MAX];      // protein residue codes
	__shared__ int8_t smat[22 * 22];
	const int t = threadIdx.x, c = blockIdx.x, lane = t & 63;
	const nsh_prep::PrepProb P = a.pp[c];
	const uint8_t *praw = a.as + P.as_off, *wraw = (P.from_genome? a.pc->genome : a.ns) + P.ns_off;
	const int al = P.al, nl = P.nl;
	if (t < 256) nt4[t] = a.pc->nt4[t], aa20[t] = a.pc->aa20[t];
	if (t < 64) codon[t] = a.pc->codon[t];
	for (int i = t; i < 22 * 22; i += NT) smat[i] = a.mat[i];
	const int ncod = nl / 3, ncodf[3] = { nl / 3, (nl - 1) / 3 > 0? (nl - 1) / 3 : 0, (nl - 2) / 3 > 0? (nl - 2) / 3 : 0 };
	const int nb = (al + ncod) / 8 + 2;
	for (int i = t; i < KMAX; i += NT) first[i] = INT_MAX;
	for (int i = t; i < 3 * ((NBIN + 1) / 2); i += NT) (&hist[0][0])[i] = 0;
	for (int i = t; i < AL_MAX / 32; i += NT) qhit[i] = 0;
	for (int i = t; i < NL_MAX / 3 / 32 + 1; i += NT) thit[i] = 0;
	if (t < 96) cnt[t] = 0;
	if (t < 2) dsum[t] = 0;
	__syncthreads();
	// one coalesced pass over each string: protein codes and 3-mer index, window codon codes and
	// base counts (per-thread counters, one LDS update per wave at the end)
	// four strides per iteration with every load issued before the first use (the loops are
	// bound by memory latency, not bandwidth)
	for (int i0 = t; i0 < al; i0 += 4 * NT) {
		int raw[4];
#pragma unroll
		for (int u = 0; u < 4; ++u) raw[u] = i0 + NT * u < al? praw[i0 + NT * u] : 0;
#pragma unroll
		for (int u = 0; u < 4; ++u) {
			const int i = i0 + NT * u;
			if (i < al) {
				const int x = aa20[raw[u]];
				pa[i] = (uint8_t)x;
				if (x < 20) atomicAdd(&cnt[14 + x], 1);
			}
		}
	}
	int n_gc = 0, n_gt = 0, n_ag = 0;
	for (int i0 = t; i0 < nl; i0 += 4 * NT) {
		int raw[4][3];
#pragma unroll
		for (int u = 0; u < 4; ++u)
#pragma unroll
			for (int e = 0; e < 3; ++e) raw[u][e] = i0 + NT * u + e < nl? wraw[i0 + NT * u + e] : 'N';
#pragma unroll
		for (int u = 0; u < 4; ++u) {
			const int i = i0 + NT * u;
			if (i < nl) {
				const int b0 = nt4[raw[u][0]], b1 = i + 1 < nl? nt4[raw[u][1]] : 4, b2 = i + 2 < nl? nt4[raw[u][2]] : 4;
				n_gc += b0 == 1 || b0 == 2;
				n_gt += b0 == 2 && b1 == 3;
				n_ag += b0 == 0 && b1 == 2;
				if (i + 2 < nl) cod[i] = b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 21;
			}
		}
	}
	__syncthreads();
	for (int i = t; i + 2 < al; i += NT) {
		const int x = pa[i], y = pa[i + 1], z = pa[i + 2];
		if (x < 20 && y < 20 && z < 20) atomicMin(&first[(x * 20 + y) * 20 + z], i);
	}
	__syncthreads();
	// anchors: codon j of frame f covers bases f+3j .. f+3j+2; a 3-mer spans codons j..j+2
	long long ds = 0, ds2 = 0;
	int n_anch[3] = {}, n_stop[3] = {}, n_x = 0, n_quart[4] = {};
#pragma unroll
	for (int f = 0; f < 3; ++f) {
		const int nc = ncodf[f];
		for (int j = t; j < nc; j += NT) {
			const int r0 = cod[f + 3 * j], r1 = j + 1 < nc? cod[f + 3 * j + 3] : 21, r2 = j + 2 < nc? cod[f + 3 * j + 6] : 21;
			n_stop[f] += r0 == 20;
			n_x += r0 == 21;
			if (r0 < 20 && r1 < 20 && r2 < 20) {
				const int q = first[(r0 * 20 + r1) * 20 + r2];
				if (q != INT_MAX) {
					++n_anch[f];
					atomicOr(&qhit[q >> 5], 1u << (q & 31));
					atomicOr(&thit[j >> 5], 1u << (j & 31));
					const int qt = (4 * q) / al;
					n_quart[0] += qt == 0, n_quart[1] += qt == 1, n_quart[2] += qt == 2, n_quart[3] += qt == 3;
					const int d = j - q, bin = (d + al) >> 3;
					atomicAdd(&hist[f][bin >> 1], 1u << ((bin & 1) * 16));
					ds += d, ds2 += (long long)d * d;
				}
			}
		}
	}
	{
		// per thread every count is at most 48 (16 codons in each of three frames, or 47 bases), so a
		// wave's sum fits 16 bits: two counts per 32-bit sum. Per thread |ds| < 2^19 and ds2 < 2^32,
		// summed as its 16-bit low and high parts.
		const int v[14] = { n_anch[0], n_anch[1], n_anch[2], n_stop[0], n_stop[1], n_stop[2], n_x, n_gc, n_gt, n_ag,
			n_quart[0], n_quart[1], n_quart[2], n_quart[3] };
#pragma unroll
		for (int k = 0; k < 14; k += 2) {
			const int sv = wave_sum(v[k] | v[k + 1] << 16);
			if (lane == 0) atomicAdd(&cnt[k], sv & 0xffff), atomicAdd(&cnt[k + 1], (int)((unsigned)sv >> 16));
		}
		const int s1 = wave_sum((int)ds), s2l = wave_sum((int)(ds2 & 0xffff)), s2h = wave_sum((int)(ds2 >> 16));
		if (lane == 0) {
			atomicAdd((unsigned long long*)&dsum[0], (unsigned long long)(long long)s1);
			atomicAdd((unsigned long long*)&dsum[1], (unsigned long long)((long long)s2h << 16) + (unsigned long long)s2l);
		}
	}
	__syncthreads();
	const int dom = cnt[0] >= cnt[1] && cnt[0] >= cnt[2]? 0 : (cnt[1] >= cnt[2]? 1 : 2);
	for (int j = t; j < ncodf[dom]; j += NT) {
		const int r = cod[dom + 3 * j];
		if (r < 20) atomicAdd(&cnt[34 + r], 1);
	}
	// bin statistics and the four fullest bins. A bin's key is count << 16 | (65535 - frame * nb - bin):
	// keys are distinct, and the largest key is the fullest bin with the smallest index. Round r
	// takes the largest key below round r-1's, so no bin is cleared and one barrier ends a round.
	int n3 = 0;
	unsigned vmax = 0;
	int hq = 0, ht = 0;
	for (int i = t; i < AL_MAX / 32; i += NT) hq += __popc(qhit[i]);
	for (int i = t; i < NL_MAX / 3 / 32 + 1; i += NT) ht += __popc(thit[i]);
	__shared__ unsigned redk[4][NT / 64];
	__shared__ int topbin[4], segsc[4][8];
	unsigned prev = 0xffffffffu;
	for (int r = 0; r < 4; ++r) {
		unsigned best = 0;
		for (int f = 0; f < 3; ++f)
			for (int bin = t; bin < nb; bin += NT) {
				const unsigned v = (hist[f][bin >> 1] >> ((bin & 1) * 16)) & 0xffff;
				const unsigned key = v << 16 | (unsigned)(65535 - (f * nb + bin));
				if (r == 0) n3 += v >= 3, vmax = max(vmax, v);
				if (v > 0 && key < prev) best = max(best, key);
			}
		best = wave_umax(best);
		if (lane == 0) redk[r][t >> 6] = best;
		if (r == 0) {
			const int s3 = wave_sum(n3), sq = wave_sum(hq), st = wave_sum(ht);
			const unsigned m = wave_umax(vmax);
			if (lane == 0) atomicAdd(&cnt[54], s3), atomicMax(&cnt[55], (int)m), atomicAdd(&cnt[56], sq), atomicAdd(&cnt[57], st);
		}
		__syncthreads();
		best = 0;
#pragma unroll
		for (int w = 0; w < NT / 64; ++w) best = max(best, redk[r][w]);
		if (t == 0) topbin[r] = best? 65535 - (int)(best & 0xffff) : -1;
		prev = best? best : 0;
	}
	__syncthreads();
	// best ungapped segment (max subarray sum, empty allowed) on each of a bin's 8 diagonals: every
	// thread takes one of NCH chunks of one diagonal of one bin (4 x 8 x NCH = NT), summarises it
	// (total, best prefix, best suffix, best inside) and 32 threads chain the NCH summaries
	constexpr int NCH = NT / 32;
	__shared__ int csum[4][8][NCH][4];
	{
		const int r = t / (8 * NCH), dd = (t / NCH) & 7, ch = t % NCH, tb = topbin[r];
		int tot = 0, pre = 0, suf = 0, ins = 0;
		if (tb >= 0) {
			const int f = tb / nb, d = (tb % nb) * 8 - al + dd;
			const int q0 = d < 0? -d : 0, q1 = min(al, ncodf[f] - d), len = q1 > q0? q1 - q0 : 0;
			const int sz = (len + NCH - 1) / NCH, qa = q0 + ch * sz, qb = min(q1, qa + sz);
			const int cb = f + 3 * d;   // cod[cb + 3q]: the window residue facing protein position q
			int run = 0, q = qa;
			for (; q + 4 <= qb; q += 4) {          // four positions per step, all loads first
				int cc[4], pp[4], v[4];
#pragma unroll
				for (int u = 0; u < 4; ++u) cc[u] = cod[cb + 3 * (q + u)], pp[u] = pa[q + u];
#pragma unroll
				for (int u = 0; u < 4; ++u) v[u] = smat[cc[u] * 22 + pp[u]];
#pragma unroll
				for (int u = 0; u < 4; ++u) {
					tot += v[u];
					pre = max(pre, tot);
					run = max(0, run + v[u]);
					ins = max(ins, run);
				}
			}
			for (; q < qb; ++q) {
				const int v = smat[cod[cb + 3 * q] * 22 + pa[q]];
				tot += v;
				pre = max(pre, tot);
				run = max(0, run + v);
				ins = max(ins, run);
			}
			suf = run;   // the best sum ending at the chunk's last position (clamped at zero)
		}
		csum[r][dd][ch][0] = tot, csum[r][dd][ch][1] = pre, csum[r][dd][ch][2] = suf, csum[r][dd][ch][3] = ins;
	}
	__syncthreads();
	if (t < 32) {
		const int r = t >> 3, dd = t & 7;
		int best = 0, s = 0;   // s: best sum ending at the previous chunk's end (>= 0)
		for (int ch = 0; ch < NCH; ++ch) {
			const int *cs = csum[r][dd][ch];
			best = max(best, max(cs[3], s + cs[1]));
			s = max(cs[2], s + cs[0]);
			s = max(s, 0);
		}
		segsc[r][dd] = best;
	}
	__syncthreads();
	// the 64 features, one per lane of the first wave: numerator and denominator are exact
	// integers in float, one division per lane
	if (t < 64) {
		const int f = t;
		const int nanch = cnt[0] + cnt[1] + cnt[2];
		const float fal = (float)al, fnl = (float)nl;
		int sg[4];
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			sg[r] = 0;
			for (int k = 0; k < 8; ++k) sg[r] = max(sg[r], segsc[r][k]);
		}
#pragma unroll
		for (int x = 1; x < 4; ++x)   // descending
#pragma unroll
			for (int y = x; y > 0; --y)
				if (sg[y] > sg[y - 1]) { const int tmp = sg[y]; sg[y] = sg[y - 1], sg[y - 1] = tmp; }
		const int tot = ncodf[0] + ncodf[1] + ncodf[2];
		const float nl1 = nl > 0? fnl : 1.f;
		float num = 0.f, den = fal;
		if (f == 2) num = fnl, den = 3.f * fal;
		else if (f >= 3 && f < 6) num = (float)cnt[f - 3];
		else if (f == 6) num = (float)nanch;
		else if (f >= 7 && f < 10) num = (float)cnt[f - 4], den = (float)(ncodf[f - 7] > 0? ncodf[f - 7] : 1);
		else if (f == 10) num = (float)cnt[6], den = (float)(tot > 0? tot : 1);
		else if (f == 11) num = (float)cnt[7], den = nl1;
		else if (f == 12 || f == 13) num = 16.f * (float)cnt[f - 4], den = nl1;
		else if (f == 14) num = (float)cnt[dom], den = (float)(nanch > 0? nanch : 1);
		else if (f == 15) num = (float)cnt[55];
		else if (f == 16) num = (float)cnt[54], den = 10.f;
		else if (f == 17) num = (float)cnt[56];
		else if (f == 18) num = (float)cnt[57], den = (float)(ncod > 0? ncod : 1);
		else if (f >= 20 && f < 40) num = (float)cnt[f - 6];
		else if (f >= 40 && f < 44) num = (float)sg[0] * (f == 40) + (float)sg[1] * (f == 41) + (float)sg[2] * (f == 42) + (float)sg[3] * (f == 43);
		else if (f == 44) num = (float)(sg[0] + sg[1] + sg[2] + sg[3]);
		else if (f == 45) num = (float)((sg[0] >= 20) + (sg[1] >= 20) + (sg[2] >= 20) + (sg[3] >= 20)), den = 4.f;
		else if (f >= 46 && f < 60) num = (float)cnt[f - 12], den = (float)(ncodf[dom] > 0? ncodf[dom] : 1);
		else if (f >= 60) num = (float)cnt[f - 50];
		float v = num / den;
		const float lg = log2f(1.f + (f == 0? fal : fnl));
		if (f < 2) v = lg;
		if (f == 19) {
			double sd = 0.;
			if (nanch > 0) {
				const double m = (double)dsum[0] / nanch, m2 = (double)dsum[1] / nanch;
				sd = m2 - m * m > 0.? sqrt(m2 - m * m) : 0.;
			}
			v = (float)(sd / (double)al);
		}
		a.out[(int64_t)c * NF + f] = v;
	}
}

} // namespace

void nsh_feat::launch(const nsh_prep::PrepProb *pp, int32_t n, const uint8_t *ns, const uint8_t *as,
		const nsh_prep::PrepConst *pc, const int8_t *mat, float *out, hipStream_t s)
{
	FeatArgs fa;
	fa.pp = pp, fa.ns = ns, fa.as = as, fa.pc = pc, fa.mat = mat, fa.n = n, fa.out = out;
	hipLaunchKernelGGL(feature_kernel, dim3(n), dim3(NT), 0, s, fa);
}

#define FOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("features: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

// features of n candidates from raw host strings (protein ASCII, DNA ASCII or 0..4); out: n x 64
extern "C" int nsh_features(int32_t n, const char *as, const int64_t *as_off, const int32_t *al, int64_t l_as,
		const char *ns, const int64_t *ns_off, const int32_t *nl, int64_t l_ns, float *out)
{
	if (!nsh_ensure_device()) return 0;
	if (n < 0 || (n > 0 && (!as || !as_off || !al || !ns || !ns_off || !nl || !out))) { nsh_set_error("features: bad arguments"); return 0; }
	if (n == 0) return 1;
	std::vector<nsh_prep::PrepProb> pp(n);
	for (int32_t k = 0; k < n; ++k) {
		if (al[k] < 1 || al[k] > AL_MAX || nl[k] < 0 || nl[k] > NL_MAX || as_off[k] < 0 || as_off[k] + al[k] > l_as ||
				ns_off[k] < 0 || ns_off[k] + nl[k] > l_ns) {
			nsh_set_error("features: candidate %d out of range (al 1..%d, nl 0..%d)", k, AL_MAX, NL_MAX);
			return 0;
		}
		pp[k] = {};
		pp[k].ns_off = ns_off[k], pp[k].as_off = as_off[k], pp[k].nl = nl[k], pp[k].al = al[k];
	}
	nsh_prep::PrepConst pc;
	std::memcpy(pc.nt4, ns_tab_nt4, 256), std::memcpy(pc.aa20, ns_tab_aa20, 256), std::memcpy(pc.codon, ns_tab_codon, 64);
	uint8_t *d_as = nullptr, *d_ns = nullptr;
	nsh_prep::PrepProb *d_pp = nullptr;
	nsh_prep::PrepConst *d_pc = nullptr;
	int8_t *d_mat = nullptr;
	float *d_out = nullptr;
	hipStream_t s = nullptr;
	int ok = 0;
	const auto t0 = std::chrono::steady_clock::now();
	{
		FOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		FOK(hipMalloc((void**)&d_as, (size_t)l_as + 4));
		FOK(hipMalloc((void**)&d_ns, (size_t)l_ns + 4));
		FOK(hipMalloc((void**)&d_pp, (size_t)n * sizeof(nsh_prep::PrepProb)));
		FOK(hipMalloc((void**)&d_pc, sizeof(pc)));
		FOK(hipMalloc((void**)&d_out, (size_t)n * NF * 4));
		FOK(hipMalloc((void**)&d_mat, 22 * 22));
		FOK(hipMemcpyAsync(d_mat, ns_mat_blosum62, 22 * 22, hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_as, as, (size_t)l_as, hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_ns, ns, (size_t)l_ns, hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_pp, pp.data(), (size_t)n * sizeof(nsh_prep::PrepProb), hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_pc, &pc, sizeof(pc), hipMemcpyHostToDevice, s));
		nsh_feat::launch(d_pp, n, d_ns, d_as, d_pc, d_mat, d_out, s);
		FOK(hipGetLastError());
		FOK(hipMemcpyAsync(out, d_out, (size_t)n * NF * 4, hipMemcpyDeviceToHost, s));
		FOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_as, d_ns, d_pp, d_pc, d_mat, d_out };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	if (s) (void)hipStreamDestroy(s);
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}

// features of n candidates whose windows are slices of an oriented genome in HBM (d_g + g_off[k],
// nl[k] bases) and whose proteins are host strings; out: n x 64 on the host
int nsh_feat::from_genome(const uint8_t *d_g, int32_t n, const int64_t *g_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, int64_t l_as, float *out)
{
	if (!nsh_ensure_device()) return 0;
	if (n < 0 || (n > 0 && (!d_g || !g_off || !nl || !as || !as_off || !al || !out))) { nsh_set_error("features: bad arguments"); return 0; }
	if (n == 0) return 1;
	std::vector<nsh_prep::PrepProb> pp(n);
	for (int32_t k = 0; k < n; ++k) {
		if (al[k] < 1 || al[k] > AL_MAX || nl[k] < 0 || nl[k] > NL_MAX || as_off[k] < 0 || as_off[k] + al[k] > l_as || g_off[k] < 0) {
			nsh_set_error("features: candidate %d out of range (al 1..%d, nl 0..%d)", k, AL_MAX, NL_MAX);
			return 0;
		}
		pp[k] = {};
		pp[k].ns_off = g_off[k], pp[k].from_genome = 1, pp[k].as_off = as_off[k], pp[k].nl = nl[k], pp[k].al = al[k];
	}
	nsh_prep::PrepConst pc;
	std::memcpy(pc.nt4, ns_tab_nt4, 256), std::memcpy(pc.aa20, ns_tab_aa20, 256), std::memcpy(pc.codon, ns_tab_codon, 64);
	pc.genome = d_g;
	uint8_t *d_as = nullptr;
	nsh_prep::PrepProb *d_pp = nullptr;
	nsh_prep::PrepConst *d_pc = nullptr;
	int8_t *d_mat = nullptr;
	float *d_out = nullptr;
	hipStream_t s = nullptr;
	int ok = 0;
	{
		FOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		FOK(hipMalloc((void**)&d_as, (size_t)l_as + 4));
		FOK(hipMalloc((void**)&d_pp, (size_t)n * sizeof(nsh_prep::PrepProb)));
		FOK(hipMalloc((void**)&d_pc, sizeof(pc)));
		FOK(hipMalloc((void**)&d_out, (size_t)n * NF * 4));
		FOK(hipMalloc((void**)&d_mat, 22 * 22));
		FOK(hipMemcpyAsync(d_mat, ns_mat_blosum62, 22 * 22, hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_as, as, (size_t)l_as, hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_pp, pp.data(), (size_t)n * sizeof(nsh_prep::PrepProb), hipMemcpyHostToDevice, s));
		FOK(hipMemcpyAsync(d_pc, &pc, sizeof(pc), hipMemcpyHostToDevice, s));
		nsh_feat::launch(d_pp, n, nullptr, d_as, d_pc, d_mat, d_out, s);
		FOK(hipGetLastError());
		FOK(hipMemcpyAsync(out, d_out, (size_t)n * NF * 4, hipMemcpyDeviceToHost, s));
		FOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_as, d_pp, d_pc, d_mat, d_out };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	if (s) (void)hipStreamDestroy(s);
	return ok;
}
