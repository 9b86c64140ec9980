// Chain-candidate features for the learned prescreen (no reference counterpart).
//
// One work-group of 256 threads per candidate (protein aa codes, DNA window as base codes).
// Anchors are exact amino-acid 3-mer matches between the protein and the three forward
// translation frames of the window; the protein's 3-mers are indexed in LDS by code
// (20^3 = 8000 entries, first position), so each window codon is one LDS lookup. Counts go to
// LDS with atomics; one thread turns them into the 64 features (layout below). Every count is
// an integer, every feature one float expression of them, so a CPU restatement reproduces the
// features exactly (up to the final float rounding).
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

__global__ void __launch_bounds__(256) feature_kernel(FeatArgs a)
{
	__shared__ int first[KMAX];
	__shared__ unsigned hist[3][(NBIN + 1) / 2];   // two 16-bit bins per word
	__shared__ unsigned qhit[AL_MAX / 32], thit[NL_MAX / 3 / 32 + 1];
	__shared__ int cnt[96];   // 0-2 anchors/frame, 3-5 stops/frame, 6 X codons, 7 GC, 8 GT, 9 AG, 10-13 quarters,
	                          // 14-33 protein comp, 34-53 dominant comp, 54 bins>=3, 55 max bin
	__shared__ long long dsum[2];
	__shared__ uint8_t nt4[256], aa20[256], codon[64];
	const int t = threadIdx.x, c = blockIdx.x;
	const nsh_prep::PrepProb P = a.pp[c];
	const uint8_t *praw = a.as + P.as_off, *wraw = a.ns + P.ns_off;
	const int al = P.al, nl = P.nl;
	nt4[t] = a.pc->nt4[t], aa20[t] = a.pc->aa20[t];
	if (t < 64) codon[t] = a.pc->codon[t];
	__syncthreads();
#define p(i) aa20[praw[i]]
#define w(i) nt4[wraw[i]]
	const int ncod = nl / 3, ncodf[3] = { nl / 3, (nl - 1) / 3 > 0? (nl - 1) / 3 : 0, (nl - 2) / 3 > 0? (nl - 2) / 3 : 0 };
	const int nb = (al + ncod) / 8 + 2;
	for (int i = t; i < KMAX; i += 256) first[i] = INT_MAX;
	for (int i = t; i < 3 * ((NBIN + 1) / 2); i += 256) (&hist[0][0])[i] = 0;
	for (int i = t; i < AL_MAX / 32; i += 256) qhit[i] = 0;
	for (int i = t; i < NL_MAX / 3 / 32 + 1; i += 256) thit[i] = 0;
	if (t < 96) cnt[t] = 0;
	if (t < 2) dsum[t] = 0;
	__syncthreads();
	for (int i = t; i + 2 < al; i += 256) {
		const int x = p(i), y = p(i + 1), z = p(i + 2);
		if (x < 20 && y < 20 && z < 20) atomicMin(&first[(x * 20 + y) * 20 + z], i);
	}
	for (int i = t; i < al; i += 256) if (p(i) < 20) atomicAdd(&cnt[14 + p(i)], 1);
	for (int i = t; i < nl; i += 256) {
		const int b = w(i);
		if (b == 1 || b == 2) atomicAdd(&cnt[7], 1);
		if (i + 1 < nl) {
			if (b == 2 && w(i + 1) == 3) atomicAdd(&cnt[8], 1);
			if (b == 0 && w(i + 1) == 2) atomicAdd(&cnt[9], 1);
		}
	}
	__syncthreads();
	// anchors: codon j of frame f covers bases f+3j .. f+3j+2; a 3-mer spans codons j..j+2
	long long ds = 0, ds2 = 0;
	for (int f = 0; f < 3; ++f) {
		for (int j = t; j < ncodf[f]; j += 256) {
			int r[3];
#pragma unroll
			for (int e = 0; e < 3; ++e) {
				const int jj = j + e;
				if (jj >= ncodf[f]) { r[e] = 21; continue; }
				const int o = f + 3 * jj, b0 = w(o), b1 = w(o + 1), b2 = w(o + 2);
				r[e] = b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 21;
			}
			if (r[0] == 20) atomicAdd(&cnt[3 + f], 1);
			if (r[0] == 21) atomicAdd(&cnt[6], 1);
			if (r[0] < 20 && r[1] < 20 && r[2] < 20) {
				const int q = first[(r[0] * 20 + r[1]) * 20 + r[2]];
				if (q != INT_MAX) {
					atomicAdd(&cnt[f], 1);
					atomicOr(&qhit[q >> 5], 1u << (q & 31));
					atomicOr(&thit[j >> 5], 1u << (j & 31));
					atomicAdd(&cnt[10 + (4 * q) / al], 1);
					const int d = j - q, bin = (d + al) >> 3;
					atomicAdd(&hist[f][bin >> 1], 1u << ((bin & 1) * 16));
					ds += d, ds2 += (long long)d * d;
				}
			}
		}
	}
	atomicAdd((unsigned long long*)&dsum[0], (unsigned long long)ds);
	atomicAdd((unsigned long long*)&dsum[1], (unsigned long long)ds2);
	__syncthreads();
	const int dom = cnt[0] >= cnt[1] && cnt[0] >= cnt[2]? 0 : (cnt[1] >= cnt[2]? 1 : 2);
	for (int j = t; j < ncodf[dom]; j += 256) {
		const int o = dom + 3 * j, b0 = w(o), b1 = w(o + 1), b2 = w(o + 2);
		const int r = b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 21;
		if (r < 20) atomicAdd(&cnt[34 + r], 1);
	}
	for (int i = t; i < 3 * nb; i += 256) {
		const int f = i / nb, bin = i % nb;
		const int v = (hist[f][bin >> 1] >> ((bin & 1) * 16)) & 0xffff;
		if (v >= 3) atomicAdd(&cnt[54], 1);
		atomicMax(&cnt[55], v);
	}
	int hq = 0, ht = 0;
	for (int i = t; i < AL_MAX / 32; i += 256) hq += __popc(qhit[i]);
	for (int i = t; i < NL_MAX / 3 / 32 + 1; i += 256) ht += __popc(thit[i]);
	__syncthreads();
	atomicAdd(&cnt[56], hq);
	atomicAdd(&cnt[57], ht);
	// the four fullest diagonal bins (ties: smaller frame * nb + bin), one at a time
	__shared__ int topbin[4], redv[4], redi[4], segsc[4][8];
	for (int r = 0; r < 4; ++r) {
		int bv = 0, bi = INT_MAX;
		for (int i = t; i < 3 * nb; i += 256) {
			const int f = i / nb, bin = i % nb;
			const int v = (hist[f][bin >> 1] >> ((bin & 1) * 16)) & 0xffff;
			if (v > bv) bv = v, bi = i;
		}
#pragma unroll
		for (int d = 32; d >= 1; d >>= 1) {
			const int ov = __shfl_xor(bv, d), oi = __shfl_xor(bi, d);
			if (ov > bv || (ov == bv && oi < bi)) bv = ov, bi = oi;
		}
		if ((t & 63) == 0) redv[t >> 6] = bv, redi[t >> 6] = bi;
		__syncthreads();
		if (t == 0) {
			int v = 0, ix = INT_MAX;
			for (int k = 0; k < 4; ++k)
				if (redv[k] > v || (redv[k] == v && redi[k] < ix)) v = redv[k], ix = redi[k];
			topbin[r] = v > 0? ix : -1;
			if (v > 0) {
				const int f = ix / nb, bin = ix % nb;
				hist[f][bin >> 1] &= ~(0xffffu << ((bin & 1) * 16));
			}
		}
		__syncthreads();
	}
	// best ungapped segment (max subarray sum, empty allowed) on each of a bin's 8 diagonals: every
	// thread takes one eighth of one diagonal of one bin (4 x 8 x 8 = 256), summarises its chunk
	// (total, best prefix, best suffix, best inside) and 32 threads chain the eight summaries
	__shared__ int csum[4][8][8][4];
	{
		const int r = t >> 6, dd = (t >> 3) & 7, ch = t & 7, tb = topbin[r];
		int tot = 0, pre = 0, suf = 0, ins = 0;
		if (tb >= 0) {
			const int f = tb / nb, d = (tb % nb) * 8 - al + dd;
			const int q0 = d < 0? -d : 0, q1 = min(al, ncodf[f] - d), len = q1 > q0? q1 - q0 : 0;
			const int sz = (len + 7) / 8, qa = q0 + ch * sz, qb = min(q1, qa + sz);
			int run = 0;
			for (int q = qa; q < qb; ++q) {
				const int o = f + 3 * (q + d), b0 = w(o), b1 = w(o + 1), b2 = w(o + 2);
				const int rr = b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 21;
				const int v = a.mat[rr * 22 + p(q)];
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
		for (int ch = 0; ch < 8; ++ch) {
			const int *cs = csum[r][dd][ch];
			best = max(best, max(cs[3], s + cs[1]));
			s = max(cs[2], s + cs[0]);
			s = max(s, 0);
		}
		segsc[r][dd] = best;
	}
	__syncthreads();
	if (t == 0) {
		float *o = a.out + (int64_t)c * NF;
		const int nanch = cnt[0] + cnt[1] + cnt[2];
		const float fal = (float)al, fnl = (float)nl;
		o[0] = log2f(1.f + fal), o[1] = log2f(1.f + fnl), o[2] = fnl / (3.f * fal);
		for (int f = 0; f < 3; ++f) o[3 + f] = (float)cnt[f] / fal;
		o[6] = (float)nanch / fal;
		for (int f = 0; f < 3; ++f) o[7 + f] = (float)cnt[3 + f] / (float)(ncodf[f] > 0? ncodf[f] : 1);
		const int tot = ncodf[0] + ncodf[1] + ncodf[2];
		o[10] = (float)cnt[6] / (float)(tot > 0? tot : 1);
		o[11] = (float)cnt[7] / (nl > 0? fnl : 1.f);
		o[12] = 16.f * (float)cnt[8] / (nl > 0? fnl : 1.f);
		o[13] = 16.f * (float)cnt[9] / (nl > 0? fnl : 1.f);
		o[14] = (float)cnt[dom] / (float)(nanch > 0? nanch : 1);
		o[15] = (float)cnt[55] / fal;
		o[16] = (float)cnt[54] / 10.f;
		o[17] = (float)cnt[56] / fal;
		o[18] = (float)cnt[57] / (float)(ncod > 0? ncod : 1);
		double sd = 0.;
		if (nanch > 0) {
			const double m = (double)dsum[0] / nanch, m2 = (double)dsum[1] / nanch;
			sd = m2 - m * m > 0.? sqrt(m2 - m * m) : 0.;
		}
		o[19] = (float)(sd / (double)al);
		for (int r = 0; r < 20; ++r) o[20 + r] = (float)cnt[14 + r] / fal;
		int sg[4];
		for (int r = 0; r < 4; ++r) {
			sg[r] = 0;
			for (int k = 0; k < 8; ++k) sg[r] = max(sg[r], segsc[r][k]);
		}
		for (int x = 1; x < 4; ++x)   // descending
			for (int y = x; y > 0 && sg[y] > sg[y - 1]; --y) { const int tmp = sg[y]; sg[y] = sg[y - 1], sg[y - 1] = tmp; }
		for (int r = 0; r < 4; ++r) o[40 + r] = (float)sg[r] / fal;
		o[44] = (float)(sg[0] + sg[1] + sg[2] + sg[3]) / fal;
		o[45] = (float)((sg[0] >= 20) + (sg[1] >= 20) + (sg[2] >= 20) + (sg[3] >= 20)) / 4.f;
		for (int r = 0; r < 14; ++r) o[46 + r] = (float)cnt[34 + r] / (float)(ncodf[dom] > 0? ncodf[dom] : 1);
		for (int k = 0; k < 4; ++k) o[60 + k] = (float)cnt[10 + k] / fal;
	}
#undef p
#undef w
}

} // namespace

void nsh_feat::launch(const nsh_prep::PrepProb *pp, int32_t n, const uint8_t *ns, const uint8_t *as,
		const nsh_prep::PrepConst *pc, const int8_t *mat, float *out, hipStream_t s)
{
	FeatArgs fa;
	fa.pp = pp, fa.ns = ns, fa.as = as, fa.pc = pc, fa.mat = mat, fa.n = n, fa.out = out;
	hipLaunchKernelGGL(feature_kernel, dim3(n), dim3(256), 0, s, fa);
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
