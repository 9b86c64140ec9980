// The refinement step of the map path on the device, for every candidate region at once
// (mp_refine_reg, map.c:141-203; the reference runs it per region on the host).
//
// A region's window (up to max_ext bases either side of it) is sketched with the short k-mers of
// the refinement (kmer2 residues, mod_bit 0) over its clean ORFs, as mp_sketch_nt4 does, and the
// window k-mers are paired with the query's k-mers of the same hash, all against all, unless a
// bucket's n1 x n2 exceeds max_ava. The pairs go to mp_chain.
//
// Only window k-mers whose hash the query also has can take part in a pair, so the window is never
// sketched as a whole: each work-group holds a bit filter of its query's hashes in LDS, sweeps its
// window in tiles staged through LDS (bases, then the codon code ending at each base), and keeps a
// position only when its k-mer's hash passes the filter and is found among the query's sorted
// k-mers, and its ORF (the run of sense codons around it, cut at the window's ends) is long enough. That is
// exactly the subset of the reference's sketch that its bucket loop can pair. The windows are
// slices of the oriented genome resident in HBM (forward strand, then the reverse complement).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int kFilterBits = 16;                          // LDS filter of the query's hashes: 2^16 bits (8 KB)
// window bases per sweep step; bases staged before and after it: the k-mer's history (at most 20
// bases) and an ORF walk of 30 codons either way
constexpr int kTile = 4096, kHalo = 120, kAhead = 96;

struct RConst { uint8_t codon[64], codon13[64]; };

__device__ inline uint32_t hash32_mask(uint32_t key, uint32_t mask)   // sketch.c:7-16
{
	key = (key + ~(key << 15)) & mask;
	key ^= key >> 10;
	key = (key + (key << 3)) & mask;
	key ^= key >> 6;
	key = (key + ~(key << 11)) & mask;
	key ^= key >> 16;
	return key;
}

// The query k-mers of the refinement (mp_sketch_prot, sketch.c:18-38: every position that ends
// kmer residues of the 13-letter alphabet, (hash >> mod_bit) << 32 | position when the hash's low
// mod_bit bits are zero), one work-group per query, in position order: with out == nullptr it
// counts them into cnt[q], else it writes them from out[out_off[q]]
__global__ void __launch_bounds__(256) qsketch_kernel(const uint8_t *aa, const int64_t *q_off, const int32_t *qlen,
		const uint8_t *aa13, int32_t kmer, int32_t mod_bit, const int64_t *out_off, int64_t *cnt, uint64_t *out)
{
	__shared__ int32_t wsum[4];
	const int q = blockIdx.x, t = threadIdx.x, L = t & 63, w = t >> 6;
	const uint8_t *sq = aa + q_off[q];
	const int32_t len = qlen[q];
	const uint32_t mask_k = (1u << kmer * 4) - 1, mask_mod = (1u << mod_bit) - 1;
	const int64_t base = out? out_off[q] : 0;
	int64_t total = 0;
	for (int32_t i0 = 0; i0 < len; i0 += 256) {
		const int32_t i = i0 + t;
		bool v = false;
		uint64_t e = 0;
		if (i < len && i >= kmer - 1) {
			uint32_t x = 0;
			v = true;
			for (int j = kmer - 1; j >= 0; --j) {   // residues i-kmer+1 .. i, the oldest in the high bits
				const uint32_t c = aa13[sq[i - j]];
				if (c >= 14) { v = false; break; }
				x = x << 4 | c;
			}
			if (v) {
				const uint32_t y = hash32_mask(x & mask_k, mask_k);
				v = (y & mask_mod) == 0;
				e = (uint64_t)(y >> mod_bit) << 32 | (uint32_t)i;
			}
		}
		const unsigned long long b = __ballot(v);
		if (L == 0) wsum[w] = __popcll(b);
		__syncthreads();
		int64_t before = total;
		for (int u = 0; u < w; ++u) before += wsum[u];
		if (out && v) out[base + before + __popcll(b & ((1ull << L) - 1))] = e;
		total += wsum[0] + wsum[1] + wsum[2] + wsum[3];
		__syncthreads();
	}
	if (!out && t == 0) cnt[q] = total;
}

// the sense codon ending at window position e (its three bases valid and not a stop)
__device__ inline bool sense_at(const uint8_t *w, int64_t e, const uint8_t *codon)
{
	const uint8_t b0 = w[e - 2], b1 = w[e - 1], b2 = w[e];
	return b0 < 4 && b1 < 4 && b2 < 4 && codon[b0 << 4 | b1 << 2 | b2] < 20;
}

// Window k-mers of each job that its query shares, written unsorted to out[cap_off[j] ..) (at most
// cap[j] of them; cnt[j] receives how many there are, so a job with cnt[j] > cap[j] is run again
// with more room) as hash << 32 | position of the k-mer's last base (the entries mp_sketch_nt4
// makes with bbit 0). One work-group of 256 per job (idx: the jobs to run, or null for all).
__global__ void __launch_bounds__(256) window_hits_kernel(const nsh_refine::Job *jobs, const int32_t *idx,
		const uint8_t *g, const int64_t *qoff, const uint64_t *qsk, int32_t kmer, int32_t min_aa_len, const RConst *rcg,
		int64_t *cnt, const int64_t *cap_off, const int64_t *cap, uint64_t *out)
{
	__shared__ uint32_t filter[(1 << kFilterBits) / 32];   // bit (h mod 2^16) set for each query hash h
	__shared__ uint8_t base[kTile + kHalo + kAhead + 2];   // window bases [t0 - kHalo - 2, t0 + kTile + kAhead)
	__shared__ uint8_t code[kTile + kHalo + kAhead];       // codon13 of the sense codon ending at t0 - kHalo + k, 0xff if none
	__shared__ uint8_t codon[64], codon13[64];
	__shared__ int32_t n_hit;
	const int32_t jb = idx? idx[blockIdx.x] : (int32_t)blockIdx.x;
	const nsh_refine::Job J = jobs[jb];
	const uint64_t *qs = qsk + qoff[J.q];
	const int64_t nq = qoff[J.q + 1] - qoff[J.q];
	const uint32_t mask_k = (uint32_t)(((uint64_t)1 << (4 * kmer)) - 1);
	const uint8_t *w = g + J.g_off;
	const int64_t len = J.len, base_out = cap_off[jb], room = cap[jb];
	if (threadIdx.x < 64) codon[threadIdx.x] = rcg->codon[threadIdx.x], codon13[threadIdx.x] = rcg->codon13[threadIdx.x];
	if (threadIdx.x == 0) n_hit = 0;
	for (int k = threadIdx.x; k < (1 << kFilterBits) / 32; k += blockDim.x) filter[k] = 0;
	__syncthreads();
	for (int64_t k = threadIdx.x; k < nq; k += blockDim.x) {
		const uint32_t b = (uint32_t)(qs[k] >> 32) & ((1u << kFilterBits) - 1);
		atomicOr(&filter[b >> 5], 1u << (b & 31));
	}
	__syncthreads();
	for (int64_t t0 = 0; t0 < len; t0 += kTile) {
		for (int k = threadIdx.x; k < kTile + kHalo + kAhead + 2; k += blockDim.x) {
			const int64_t p = t0 - kHalo - 2 + k;
			base[k] = p >= 0 && p < len? w[p] : 4;
		}
		__syncthreads();
		for (int k = threadIdx.x; k < kTile + kHalo + kAhead; k += blockDim.x) {   // position t0 - kHalo + k ends a codon?
			const uint8_t b0 = base[k], b1 = base[k + 1], b2 = base[k + 2];
			const int cd = b0 << 4 | b1 << 2 | b2;
			code[k] = b0 < 4 && b1 < 4 && b2 < 4 && codon[cd] < 20? codon13[cd] : 0xff;
		}
		__syncthreads();
		for (int k = threadIdx.x; k < kTile; k += blockDim.x) {
			const int64_t i = t0 + k;   // the k-mer's last base
			if (i >= len || i - 3 * (kmer - 1) - 2 < 0) continue;
			uint32_t x = 0;
			bool ok = true;
			for (int c = kmer - 1; c >= 0 && ok; --c) {
				const uint8_t v = code[kHalo + k - 3 * c];
				ok = v != 0xff;
				x = (x << 4 | v) & mask_k;
			}
			if (!ok) continue;
			const uint32_t y = hash32_mask(x, mask_k), fb = y & ((1u << kFilterBits) - 1);
			if (!(filter[fb >> 5] >> (fb & 31) & 1)) continue;
			// past the filter: a binary search of the query's sorted k-mers decides
			int64_t lo = 0, hi = nq;
			while (lo < hi) {
				const int64_t m = (lo + hi) >> 1;
				if ((uint32_t)(qs[m] >> 32) < y) lo = m + 1;
				else hi = m;
			}
			if (lo >= nq || (uint32_t)(qs[lo] >> 32) != y) continue;
			// its ORF: the run of sense codons around it in its frame, cut at the window's ends, need
			// only be known to be min_aa_len codons long; staged codes serve the walk, global memory
			// beyond them
			auto sense = [&](int64_t e) {
				const int64_t c = e - (t0 - kHalo);
				return c >= 0 && c < kTile + kHalo + kAhead? code[c] != 0xff : sense_at(w, e, codon);
			};
			int64_t s = i - 3 * (kmer - 1), e = i;
			while ((e - s) / 3 + 1 < min_aa_len && s - 3 >= 2 && sense(s - 3)) s -= 3;
			while ((e - s) / 3 + 1 < min_aa_len && e + 3 < len && sense(e + 3)) e += 3;
			if ((e - s) / 3 + 1 < min_aa_len) continue;
			const int32_t slot = atomicAdd(&n_hit, 1);
			if (slot < room) out[base_out + slot] = (uint64_t)y << 32 | (uint64_t)i;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) cnt[jb] = n_hit;
}

// The bucket pairs of each job (map.c:165-198): a window entry with hash h pairs with every query
// k-mer of hash h when n1 x n2 <= max_ava (n1, n2: the bucket's window and query counts, their
// product in int32 as the reference computes it). Counted (out == nullptr) or written as
// window position << 32 | query position. went: each job's entries, sorted, at [wbeg[j], wend[j]).
__global__ void __launch_bounds__(256) pairs_kernel(const nsh_refine::Job *jobs, const int64_t *wbeg,
		const int64_t *wend, const uint64_t *went, const int64_t *qoff, const uint64_t *qsk, int32_t max_ava, int64_t *cnt,
		const int64_t *poff, uint64_t *out)
{
	__shared__ int64_t n_pair;
	const nsh_refine::Job J = jobs[blockIdx.x];
	const uint64_t *E = went + wbeg[blockIdx.x], *Q = qsk + qoff[J.q];
	const int64_t ne = wend[blockIdx.x] - wbeg[blockIdx.x], nq = qoff[J.q + 1] - qoff[J.q];
	if (threadIdx.x == 0) n_pair = 0;
	__syncthreads();
	auto lower = [](const uint64_t *a, int64_t n, uint32_t h) {
		int64_t lo = 0, hi = n;
		while (lo < hi) {
			const int64_t m = (lo + hi) >> 1;
			if ((uint32_t)(a[m] >> 32) < h) lo = m + 1;
			else hi = m;
		}
		return lo;
	};
	const int64_t pbase = poff? poff[blockIdx.x] : 0;
	for (int64_t e = threadIdx.x; e < ne; e += blockDim.x) {
		const uint32_t h = (uint32_t)(E[e] >> 32);
		const int64_t w0 = lower(E, ne, h), w1 = lower(E, ne, h + 1);
		const int64_t q0 = lower(Q, nq, h), q1 = h == 0xffffffffu? nq : lower(Q, nq, h + 1);
		const int32_t n1 = (int32_t)(w1 - w0), n2 = (int32_t)(q1 - q0);
		if (n1 <= 0 || n2 <= 0 || (int32_t)((uint32_t)n1 * (uint32_t)n2) > max_ava) continue;
		const int64_t at = atomicAdd((unsigned long long*)&n_pair, (unsigned long long)n2);
		if (out)
			for (int64_t k = q0; k < q1; ++k)
				out[pbase + at + (k - q0)] = (uint64_t)(uint32_t)E[e] << 32 | ((uint32_t)Q[k] & 0x7fffffffu);
	}
	__syncthreads();
	if (!out && threadIdx.x == 0) cnt[blockIdx.x] = n_pair;
}

// mp_refine_reg's choice among a job's chains (map.c:240-247 without the --accuracy rules): the
// first of the highest score; its anchors are copied to a compact array at an offset reserved with
// one atomic. best[j] = { offset, the chain's score << 32 | count } (count 0: no chain)
__global__ void __launch_bounds__(64) best_chain_kernel(const int64_t *poff, const int32_t *nu, const uint64_t *u,
		const uint64_t *ao, unsigned long long *total, int64_t *best, uint64_t *compact)
{
	__shared__ int64_t k_first, off;
	__shared__ int32_t n_a;
	const int32_t j = blockIdx.x;
	const int32_t n_u = nu[j];
	const uint64_t *uj = u + poff[j];
	if (threadIdx.x == 0) {
		int32_t max_i = 0;
		int64_t k = 0;
		if (n_u > 0) {
			uint64_t max_sc = uj[0] >> 32;
			for (int32_t i = 1; i < n_u; ++i)
				if (max_sc < uj[i] >> 32) max_sc = uj[i] >> 32, max_i = i;
			for (int32_t i = 0; i < max_i; ++i) k += (uint32_t)uj[i];
		}
		n_a = n_u > 0? (int32_t)(uint32_t)uj[max_i] : 0;
		k_first = k;
		off = n_a > 0? (int64_t)atomicAdd(total, (unsigned long long)n_a) : 0;
		best[2 * j] = off, best[2 * j + 1] = n_u > 0? (int64_t)uj[max_i] : 0;
	}
	__syncthreads();
	const uint64_t *src = ao + poff[j] + k_first;
	for (int32_t t = threadIdx.x; t < n_a; t += blockDim.x) compact[off + t] = src[t];
}

__global__ void revcomp_contigs_kernel(const uint8_t *fwd, const int64_t *off, const int64_t *len, int32_t n_ctg,
		uint8_t *rc)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int32_t c = blockIdx.y;
	if (c >= n_ctg || k >= len[c]) return;
	const uint8_t b = fwd[off[c] + len[c] - 1 - k];
	rc[off[c] + k] = b >= 4? b : (uint8_t)(3 - b);
}

#define RFOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("refine: %s: %s", #x, hipGetErrorString(e_)); goto done; } } while (0)

} // namespace

int nsh_refine::oriented_genome(const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len, int32_t n_ctg,
		int64_t total, uint8_t **d_g)
{
	*d_g = nullptr;
	if (!nsh_ensure_device()) return 0;
	int ok = 0;
	uint8_t *g = nullptr;
	int64_t *d_off = nullptr, *d_len = nullptr, max_len = 1;
	hipStream_t s = nullptr;
	for (int32_t c = 0; c < n_ctg; ++c) max_len = std::max(max_len, ctg_len[c]);
	{
		RFOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		RFOK(hipMalloc((void**)&g, (size_t)(2 * total + 64)));
		RFOK(hipMalloc((void**)&d_off, (size_t)n_ctg * 8));
		RFOK(hipMalloc((void**)&d_len, (size_t)n_ctg * 8));
		if (!nsh_copy_to_device(g, nt4, (size_t)total, s)) goto done;
		RFOK(hipMemcpyAsync(d_off, ctg_off, (size_t)n_ctg * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMemcpyAsync(d_len, ctg_len, (size_t)n_ctg * 8, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(revcomp_contigs_kernel, dim3((unsigned)((max_len + 255) / 256), (unsigned)n_ctg), dim3(256), 0,
				s, g, d_off, d_len, n_ctg, g + total);
		RFOK(hipGetLastError());
		RFOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		ok = 1;
	}
done:
	if (s) (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s);
	if (d_off) (void)hipFree(d_off);
	if (d_len) (void)hipFree(d_len);
	if (ok) *d_g = g;
	else if (g) (void)hipFree(g);
	return ok;
}

int nsh_refine::run(const uint8_t *d_g, const std::vector<Job> &jobs, const char *qs, const int64_t *q_off,
		const int32_t *qlen, int32_t n_q, int64_t n_res, int32_t kmer, int32_t min_aa_len, int32_t max_ava,
		const nsh_chain::Params &cp, std::vector<uint64_t> &best_u, std::vector<int64_t> &best_off,
		std::vector<uint64_t> &anchors, int64_t *n_pairs)
{
	const int32_t n = (int32_t)jobs.size();
	std::vector<int64_t> p_off((size_t)n + 1, 0);
	best_u.assign(n, 0), best_off.assign(n, 0), anchors.clear();
	if (n_pairs) *n_pairs = 0;
	if (n == 0) return 1;
	if (kmer < 1 || kmer > 7) { nsh_set_error("refine: k-mer length %d outside 1..7", kmer); return 0; }
	if (!nsh_ensure_device()) return 0;
	nsh_step_clock clk("refine");
	int ok = 0;
	hipStream_t s = nullptr;
	Job *d_jobs = nullptr;
	int64_t *d_qoff = nullptr, *d_cnt = nullptr, *d_poff = nullptr;
	uint64_t *d_qsk = nullptr, *d_w = nullptr, *d_ws = nullptr, *d_p = nullptr, *d_ps = nullptr, *d_u = nullptr;
	uint64_t *d_ao = nullptr;
	int32_t *d_nu = nullptr;
	int64_t *d_best = nullptr;
	uint64_t *d_compact = nullptr;
	unsigned long long *d_total = nullptr;
	int32_t *d_over = nullptr;
	void *d_tmp = nullptr;
	size_t tmp_cap = 0;
	RConst rc;
	RConst *d_rc = nullptr;
	std::vector<int64_t> cnt(n), cap(n), cap_off(n), wend(n), pc(n);
	int64_t *d_cap = nullptr, *d_capoff = nullptr, *d_wend = nullptr;
	uint8_t *d_qs = nullptr, *d_aa13 = nullptr;
	int64_t *d_qoffaa = nullptr, *d_qcnt = nullptr;
	int32_t *d_qlen = nullptr;
	uint64_t *d_qsk_raw = nullptr;
	std::vector<int64_t> qoff((size_t)n_q + 1, 0);
	int64_t n_qsk = 0;
	for (int t = 0; t < 64; ++t) rc.codon[t] = ns_tab_codon[t], rc.codon13[t] = ns_tab_codon13[t];
	// a segmented sort of the keys in [beg[j], end[j]) of each segment (others untouched)
	auto sort_segments = [&](const uint64_t *in, uint64_t *outk, int64_t total, const int64_t *d_beg,
			const int64_t *d_end) -> bool {
		if (total <= 0) return true;
		size_t need = 0;
		if (hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, need, in, outk, (int)total, n, d_beg, d_end, 0, 64, s) !=
				hipSuccess) return false;
		if (need > tmp_cap) {
			if (d_tmp) (void)hipFree(d_tmp);
			d_tmp = nullptr, tmp_cap = 0;
			if (hipMalloc(&d_tmp, need + 256) != hipSuccess) return false;
			tmp_cap = need + 256;
		}
		return hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, need, in, outk, (int)total, n, d_beg, d_end, 0, 64, s) ==
				hipSuccess;
	};
	{
		RFOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		// the query k-mers on the device: counts, offsets, the k-mers in position order, then each
		// query's k-mers sorted (the order of the reference's per-query sort: hash, then position)
		RFOK(hipMalloc((void**)&d_qs, (size_t)n_res + 8));
		RFOK(hipMalloc((void**)&d_qoffaa, (size_t)n_q * 8 + 8));
		RFOK(hipMalloc((void**)&d_qlen, (size_t)n_q * 4 + 4));
		RFOK(hipMalloc((void**)&d_qcnt, (size_t)n_q * 8 + 8));
		RFOK(hipMalloc((void**)&d_aa13, 256));
		if (n_res > 0) RFOK(hipMemcpyAsync(d_qs, qs, (size_t)n_res, hipMemcpyHostToDevice, s));
		if (n_q > 0) {
			RFOK(hipMemcpyAsync(d_qoffaa, q_off, (size_t)n_q * 8, hipMemcpyHostToDevice, s));
			RFOK(hipMemcpyAsync(d_qlen, qlen, (size_t)n_q * 4, hipMemcpyHostToDevice, s));
		}
		RFOK(hipMemcpyAsync(d_aa13, ns_tab_aa13, 256, hipMemcpyHostToDevice, s));
		if (n_q > 0) {
			hipLaunchKernelGGL(qsketch_kernel, dim3(n_q), dim3(256), 0, s, d_qs, d_qoffaa, d_qlen, d_aa13, kmer, 0,
					(const int64_t*)nullptr, d_qcnt, (uint64_t*)nullptr);
			RFOK(hipGetLastError());
			RFOK(hipMemcpyAsync(qoff.data() + 1, d_qcnt, (size_t)n_q * 8, hipMemcpyDeviceToHost, s));
			RFOK(hipStreamSynchronize(s));
		}
		for (int32_t q = 0; q < n_q; ++q) qoff[q + 1] += qoff[q];
		n_qsk = qoff[n_q];
		RFOK(hipMalloc((void**)&d_jobs, (size_t)n * sizeof(Job)));
		RFOK(hipMalloc((void**)&d_qoff, qoff.size() * 8));
		RFOK(hipMalloc((void**)&d_qsk, ((size_t)n_qsk + 1) * 8));
		RFOK(hipMalloc((void**)&d_qsk_raw, ((size_t)n_qsk + 1) * 8));
		RFOK(hipMemcpyAsync(d_qoff, qoff.data(), qoff.size() * 8, hipMemcpyHostToDevice, s));
		if (n_qsk > 0) {
			hipLaunchKernelGGL(qsketch_kernel, dim3(n_q), dim3(256), 0, s, d_qs, d_qoffaa, d_qlen, d_aa13, kmer, 0,
					(const int64_t*)d_qoff, (int64_t*)nullptr, d_qsk_raw);
			RFOK(hipGetLastError());
			size_t need = 0;
			RFOK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, need, d_qsk_raw, d_qsk, (int)n_qsk, n_q, d_qoff,
					d_qoff + 1, 0, 64, s));
			RFOK(hipMalloc(&d_tmp, need + 256));
			tmp_cap = need + 256;
			RFOK(hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, need, d_qsk_raw, d_qsk, (int)n_qsk, n_q, d_qoff,
					d_qoff + 1, 0, 64, s));
		}
		nsh_count_launch(3);
		clk.lap("query k-mers", n_qsk);
		RFOK(hipMalloc((void**)&d_cnt, (size_t)n * 8));
		RFOK(hipMalloc((void**)&d_cap, (size_t)n * 8));
		RFOK(hipMalloc((void**)&d_capoff, (size_t)n * 8));
		RFOK(hipMalloc((void**)&d_wend, (size_t)n * 8));
		RFOK(hipMalloc((void**)&d_poff, ((size_t)n + 1) * 8));
		RFOK(hipMalloc((void**)&d_rc, sizeof(RConst)));
		RFOK(hipMemcpyAsync(d_jobs, jobs.data(), (size_t)n * sizeof(Job), hipMemcpyHostToDevice, s));
		RFOK(hipMemcpyAsync(d_rc, &rc, sizeof(RConst), hipMemcpyHostToDevice, s));
		// the shared window k-mers of every job, in one sweep: each job writes at most its room
		// (its window length, at most 2,048, a few hundred being usual); a job with more is swept
		// again with room for all of them
		for (int32_t j = 0; j < n; ++j) cap[j] = std::min<int64_t>(std::max<int32_t>(jobs[j].len, 1), 2048);
		int64_t n_cap = 0;
		for (int32_t j = 0; j < n; ++j) cap_off[j] = n_cap, n_cap += cap[j];
		if (n_cap > INT32_MAX / 2) { nsh_set_error("refine: %lld window k-mer slots exceed the sort's range", (long long)n_cap); goto done; }
		RFOK(hipMalloc((void**)&d_w, (size_t)(n_cap + 1) * 8));
		RFOK(hipMemcpyAsync(d_cap, cap.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMemcpyAsync(d_capoff, cap_off.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(window_hits_kernel, dim3(n), dim3(256), 0, s, d_jobs, (const int32_t*)nullptr, d_g, d_qoff,
				d_qsk, kmer, min_aa_len, d_rc, d_cnt, (const int64_t*)d_capoff, (const int64_t*)d_cap, d_w);
		RFOK(hipGetLastError());
		RFOK(hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
		RFOK(hipStreamSynchronize(s));
		nsh_count_launch(1);
		{   // jobs with more than their room: swept again into room appended after everything else
			std::vector<int32_t> over;
			int64_t extra = 0;
			for (int32_t j = 0; j < n; ++j)
				if (cnt[j] > cap[j]) over.push_back(j), cap_off[j] = n_cap + extra, cap[j] = cnt[j], extra += cnt[j];
			if (!over.empty()) {
				if (n_cap + extra > INT32_MAX / 2) { nsh_set_error("refine: too many window k-mers"); goto done; }
				uint64_t *w2 = nullptr;
				RFOK(hipMalloc((void**)&w2, (size_t)(n_cap + extra + 1) * 8));
				if (hipMemcpyAsync(w2, d_w, (size_t)n_cap * 8, hipMemcpyDeviceToDevice, s) != hipSuccess) {
					(void)hipFree(w2);
					nsh_set_error("refine: copying the window k-mers failed");
					goto done;
				}
				(void)hipStreamSynchronize(s);
				(void)hipFree(d_w);
				d_w = w2, n_cap += extra;
				RFOK(hipMalloc((void**)&d_over, over.size() * 4));
				RFOK(hipMemcpyAsync(d_over, over.data(), over.size() * 4, hipMemcpyHostToDevice, s));
				RFOK(hipMemcpyAsync(d_cap, cap.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
				RFOK(hipMemcpyAsync(d_capoff, cap_off.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
				hipLaunchKernelGGL(window_hits_kernel, dim3((unsigned)over.size()), dim3(256), 0, s, d_jobs,
						(const int32_t*)d_over, d_g, d_qoff, d_qsk, kmer, min_aa_len, d_rc, d_cnt, (const int64_t*)d_capoff,
						(const int64_t*)d_cap, d_w);
				RFOK(hipGetLastError());
				RFOK(hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
				RFOK(hipStreamSynchronize(s));
				nsh_count_launch(1);
				for (int32_t j : over)
					if (cnt[j] != cap[j]) { nsh_set_error("refine: window k-mer counts changed between sweeps"); goto done; }
			}
		}
		clk.lap("window k-mers", n);
		for (int32_t j = 0; j < n; ++j) wend[j] = cap_off[j] + cnt[j];
		RFOK(hipMemcpyAsync(d_wend, wend.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMalloc((void**)&d_ws, (size_t)(n_cap + 1) * 8));
		if (!sort_segments(d_w, d_ws, n_cap, d_capoff, d_wend)) { nsh_set_error("refine: sorting the window k-mers failed"); goto done; }
		RFOK(hipStreamSynchronize(s));
		clk.lap("sort window k-mers");
		// the bucket pairs: count, offsets, write, sort per job
		hipLaunchKernelGGL(pairs_kernel, dim3(n), dim3(256), 0, s, d_jobs, (const int64_t*)d_capoff,
				(const int64_t*)d_wend, d_ws, d_qoff, d_qsk, max_ava, d_cnt, (const int64_t*)nullptr, (uint64_t*)nullptr);
		RFOK(hipGetLastError());
		RFOK(hipMemcpyAsync(pc.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
		RFOK(hipStreamSynchronize(s));
		for (int32_t j = 0; j < n; ++j) p_off[j + 1] = p_off[j] + pc[j];
		const int64_t n_p = p_off[n];
		if (n_pairs) *n_pairs = n_p;
		RFOK(hipMemcpyAsync(d_poff, p_off.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMalloc((void**)&d_p, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_ps, (size_t)(n_p + 1) * 8));
		hipLaunchKernelGGL(pairs_kernel, dim3(n), dim3(256), 0, s, d_jobs, (const int64_t*)d_capoff,
				(const int64_t*)d_wend, d_ws, d_qoff, d_qsk, max_ava, d_cnt, (const int64_t*)d_poff, d_p);
		RFOK(hipGetLastError());
		if (!sort_segments(d_p, d_ps, n_p, d_poff, d_poff + 1)) { nsh_set_error("refine: sorting the anchor pairs failed"); goto done; }
		RFOK(hipStreamSynchronize(s));
		clk.lap("pairs and their sort", p_off[n]);
		nsh_count_launch(3);
		// mp_chain of every job's pairs (map.c:203), then the chains and their anchors to the host
		RFOK(hipMalloc((void**)&d_u, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_ao, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_nu, (size_t)n * 4));
		if (n_p > 0) {
			if (!nsh_chain::run(s, n, d_poff, nullptr, pc.data(), d_ps, n_p, cp, d_nu, d_u, d_ao)) goto done;
		} else RFOK(hipMemsetAsync(d_nu, 0, (size_t)n * 4, s));
		clk.lap("chains");
		// the chosen chain of every job and its anchors, compacted on the device
		RFOK(hipMalloc((void**)&d_best, (size_t)n * 16));
		RFOK(hipMalloc((void**)&d_compact, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_total, 8));
		RFOK(hipMemsetAsync(d_total, 0, 8, s));
		hipLaunchKernelGGL(best_chain_kernel, dim3(n), dim3(64), 0, s, (const int64_t*)d_poff, (const int32_t*)d_nu,
				(const uint64_t*)d_u, (const uint64_t*)d_ao, d_total, d_best, d_compact);
		RFOK(hipGetLastError());
		nsh_count_launch(1);
		{
			unsigned long long n_best = 0;
			std::vector<int64_t> b2((size_t)2 * n);
			RFOK(hipMemcpyAsync(&n_best, d_total, 8, hipMemcpyDeviceToHost, s));
			RFOK(hipMemcpyAsync(b2.data(), d_best, (size_t)n * 16, hipMemcpyDeviceToHost, s));
			RFOK(hipStreamSynchronize(s));
			for (int32_t j = 0; j < n; ++j) best_off[j] = b2[2 * j], best_u[j] = (uint64_t)b2[2 * j + 1];
			anchors.resize((size_t)n_best);
			if (n_best && !nsh_copy_to_host(anchors.data(), d_compact, (size_t)n_best * 8, s)) goto done;
		}
		RFOK(hipStreamSynchronize(s));
		clk.lap("results to the host");
		ok = 1;
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_jobs, d_qoff, d_cnt, d_cap, d_capoff, d_wend, d_poff, d_qsk, d_w, d_ws, d_p, d_ps, d_u, d_ao, d_nu,
			d_tmp, d_rc, d_best, d_compact, d_total, d_over, d_qs, d_aa13, d_qoffaa, d_qcnt, d_qlen, d_qsk_raw };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	if (s) (void)hipStreamDestroy(s);
	return ok;
}
