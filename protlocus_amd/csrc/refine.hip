// The refinement step of the map path on the device, for every candidate region at once
// (mp_refine_reg, map.c:141-203; the reference runs it per region on the host).
//
// A region's window (up to max_ext bases either side of it) is sketched with the short k-mers of
// the refinement (kmer2 residues, mod_bit 0) over its clean ORFs, as mp_sketch_nt4 does, and the
// window k-mers are paired with the query's k-mers of the same hash, all against all, unless a
// bucket's n1 x n2 exceeds max_ava. The pairs go to mp_chain.
//
// Only window k-mers whose hash the query also has can take part in a pair, so the window is never
// sketched as a whole: each work-group holds its query's hashes in an LDS hash set, sweeps its
// window in tiles staged through LDS, and keeps a position only when its k-mer's hash is in the set
// and its ORF (the run of sense codons around it, cut at the window's ends) is long enough. That is
// exactly the subset of the reference's sketch that its bucket loop can pair. The windows are
// slices of the oriented genome resident in HBM (forward strand, then the reverse complement).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int kSetBits = 14, kSetSize = 1 << kSetBits;   // LDS hash set: 16,384 slots (64 KB)
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr int kTile = 4096, kHalo = 24;                   // window bases per sweep step; k-mer history

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

__device__ inline uint32_t slot_of(uint32_t h) { return (h * 2654435761u) >> (32 - kSetBits); }

// the sense codon ending at window position e (its three bases valid and not a stop)
__device__ inline bool sense_at(const uint8_t *w, int64_t e, const uint8_t *codon)
{
	const uint8_t b0 = w[e - 2], b1 = w[e - 1], b2 = w[e];
	return b0 < 4 && b1 < 4 && b2 < 4 && codon[b0 << 4 | b1 << 2 | b2] < 20;
}

// Window k-mers of each job that its query shares: counted per job (out == nullptr) or written to
// out[woff[j] ..) as hash << 32 | position of the k-mer's last base (the entries mp_sketch_nt4 makes
// with bbit 0). One work-group of 256 per job.
__global__ void __launch_bounds__(256) window_hits_kernel(const nsh_refine::Job *jobs, const uint8_t *g,
		const int64_t *qoff, const uint64_t *qsk, int32_t kmer, int32_t min_aa_len, const RConst *rcg, int64_t *cnt,
		const int64_t *woff, uint64_t *out)
{
	__shared__ uint32_t set[kSetSize];
	__shared__ uint8_t tile[kTile + kHalo];
	__shared__ uint8_t codon[64], codon13[64];
	__shared__ int32_t n_hit;
	const nsh_refine::Job J = jobs[blockIdx.x];
	const uint64_t *qs = qsk + qoff[J.q];
	const int64_t nq = qoff[J.q + 1] - qoff[J.q];
	const bool in_lds = nq <= kSetSize / 2;
	const uint32_t mask_k = (uint32_t)(((uint64_t)1 << (4 * kmer)) - 1);
	const uint8_t *w = g + J.g_off;
	const int64_t len = J.len;
	if (threadIdx.x < 64) codon[threadIdx.x] = rcg->codon[threadIdx.x], codon13[threadIdx.x] = rcg->codon13[threadIdx.x];
	if (threadIdx.x == 0) n_hit = 0;
	if (in_lds) {
		for (int k = threadIdx.x; k < kSetSize; k += blockDim.x) set[k] = kEmpty;
		__syncthreads();
		for (int64_t k = threadIdx.x; k < nq; k += blockDim.x) {
			const uint32_t h = (uint32_t)(qs[k] >> 32);
			for (uint32_t s = slot_of(h);; s = (s + 1) & (kSetSize - 1)) {
				const uint32_t old = atomicCAS(&set[s], kEmpty, h);
				if (old == kEmpty || old == h) break;
			}
		}
	}
	__syncthreads();
	const int64_t wbase = woff? woff[blockIdx.x] : 0;
	for (int64_t t0 = 0; t0 < len; t0 += kTile) {
		for (int k = threadIdx.x; k < kTile + kHalo; k += blockDim.x) {
			const int64_t p = t0 - kHalo + k;
			tile[k] = p >= 0 && p < len? w[p] : 4;
		}
		__syncthreads();
		for (int k = threadIdx.x; k < kTile; k += blockDim.x) {
			const int64_t i = t0 + k;   // the k-mer's last base
			if (i >= len || i - 3 * (kmer - 1) - 2 < 0) continue;
			const uint8_t *tw = tile + kHalo + k - i;   // tw[p] is window base p for p in [i - kHalo, i]
			uint32_t x = 0;
			bool ok = true;
			for (int c = kmer - 1; c >= 0 && ok; --c) {
				const int64_t e = i - 3 * c;
				const uint8_t b0 = tw[e - 2], b1 = tw[e - 1], b2 = tw[e];
				const int cd = b0 << 4 | b1 << 2 | b2;
				ok = b0 < 4 && b1 < 4 && b2 < 4 && codon[cd] < 20;
				x = (x << 4 | codon13[cd]) & mask_k;
			}
			if (!ok) continue;
			const uint32_t y = hash32_mask(x, mask_k);
			bool hit = false;
			if (in_lds) {
				for (uint32_t s = slot_of(y);; s = (s + 1) & (kSetSize - 1)) {
					const uint32_t v = set[s];
					if (v == y) { hit = true; break; }
					if (v == kEmpty) break;
				}
			} else {   // a query longer than the LDS set holds: binary search of its sorted k-mers
				int64_t lo = 0, hi = nq;
				while (lo < hi) {
					const int64_t m = (lo + hi) >> 1;
					if ((uint32_t)(qs[m] >> 32) < y) lo = m + 1;
					else hi = m;
				}
				hit = lo < nq && (uint32_t)(qs[lo] >> 32) == y;
			}
			if (!hit) continue;
			// its ORF: the run of sense codons around it in its frame, cut at the window's ends
			int64_t s = i - 3 * (kmer - 1), e = i;
			while (s - 3 >= 2 && sense_at(w, s - 3, codon)) s -= 3;
			while (e + 3 < len && sense_at(w, e + 3, codon)) e += 3;
			if ((e - s) / 3 + 1 < min_aa_len) continue;
			const int32_t slot = atomicAdd(&n_hit, 1);
			if (out) out[wbase + slot] = (uint64_t)y << 32 | (uint64_t)i;
		}
		__syncthreads();
	}
	if (!out && threadIdx.x == 0) cnt[blockIdx.x] = n_hit;
}

// The bucket pairs of each job (map.c:165-198): a window entry with hash h pairs with every query
// k-mer of hash h when n1 x n2 <= max_ava (n1, n2: the bucket's window and query counts, their
// product in int32 as the reference computes it). Counted (out == nullptr) or written as
// window position << 32 | query position. went: each job's entries sorted.
__global__ void __launch_bounds__(256) pairs_kernel(const nsh_refine::Job *jobs, const int64_t *woff,
		const uint64_t *went, const int64_t *qoff, const uint64_t *qsk, int32_t max_ava, int64_t *cnt,
		const int64_t *poff, uint64_t *out)
{
	__shared__ int64_t n_pair;
	const nsh_refine::Job J = jobs[blockIdx.x];
	const uint64_t *E = went + woff[blockIdx.x], *Q = qsk + qoff[J.q];
	const int64_t ne = woff[blockIdx.x + 1] - woff[blockIdx.x], nq = qoff[J.q + 1] - qoff[J.q];
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

int nsh_refine::run(const uint8_t *d_g, const std::vector<Job> &jobs, const std::vector<int64_t> &qoff,
		const std::vector<uint64_t> &qsk, int32_t kmer, int32_t min_aa_len, int32_t max_ava, const nsh_chain::Params &cp,
		std::vector<int64_t> &p_off, std::vector<int32_t> &n_u, std::vector<uint64_t> &u, std::vector<uint64_t> &ao)
{
	const int32_t n = (int32_t)jobs.size();
	p_off.assign((size_t)n + 1, 0), n_u.assign(n, 0), u.clear(), ao.clear();
	if (n == 0) return 1;
	if (kmer < 1 || kmer > 7) { nsh_set_error("refine: k-mer length %d outside 1..7", kmer); return 0; }
	if (!nsh_ensure_device()) return 0;
	int ok = 0;
	hipStream_t s = nullptr;
	Job *d_jobs = nullptr;
	int64_t *d_qoff = nullptr, *d_cnt = nullptr, *d_woff = nullptr, *d_poff = nullptr;
	uint64_t *d_qsk = nullptr, *d_w = nullptr, *d_ws = nullptr, *d_p = nullptr, *d_ps = nullptr, *d_u = nullptr;
	uint64_t *d_ao = nullptr;
	int32_t *d_nu = nullptr;
	void *d_tmp = nullptr;
	size_t tmp_cap = 0;
	RConst rc;
	RConst *d_rc = nullptr;
	std::vector<int64_t> cnt(n), woff((size_t)n + 1, 0), pc(n);
	for (int t = 0; t < 64; ++t) rc.codon[t] = ns_tab_codon[t], rc.codon13[t] = ns_tab_codon13[t];
	auto sort_segments = [&](const uint64_t *in, uint64_t *outk, int64_t total, const int64_t *d_seg) -> bool {
		if (total <= 0) return true;
		size_t need = 0;
		if (hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, need, in, outk, (int)total, n, d_seg, d_seg + 1, 0, 64,
				s) != hipSuccess) return false;
		if (need > tmp_cap) {
			if (d_tmp) (void)hipFree(d_tmp);
			d_tmp = nullptr, tmp_cap = 0;
			if (hipMalloc(&d_tmp, need + 256) != hipSuccess) return false;
			tmp_cap = need + 256;
		}
		return hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, need, in, outk, (int)total, n, d_seg, d_seg + 1, 0, 64,
				s) == hipSuccess;
	};
	{
		RFOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		RFOK(hipMalloc((void**)&d_jobs, (size_t)n * sizeof(Job)));
		RFOK(hipMalloc((void**)&d_qoff, qoff.size() * 8));
		RFOK(hipMalloc((void**)&d_qsk, (qsk.size() + 1) * 8));
		RFOK(hipMalloc((void**)&d_cnt, (size_t)n * 8));
		RFOK(hipMalloc((void**)&d_woff, ((size_t)n + 1) * 8));
		RFOK(hipMalloc((void**)&d_poff, ((size_t)n + 1) * 8));
		RFOK(hipMalloc((void**)&d_rc, sizeof(RConst)));
		RFOK(hipMemcpyAsync(d_jobs, jobs.data(), (size_t)n * sizeof(Job), hipMemcpyHostToDevice, s));
		RFOK(hipMemcpyAsync(d_qoff, qoff.data(), qoff.size() * 8, hipMemcpyHostToDevice, s));
		if (!qsk.empty()) RFOK(hipMemcpyAsync(d_qsk, qsk.data(), qsk.size() * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMemcpyAsync(d_rc, &rc, sizeof(RConst), hipMemcpyHostToDevice, s));
		// the shared window k-mers: count, offsets, write, sort per job
		hipLaunchKernelGGL(window_hits_kernel, dim3(n), dim3(256), 0, s, d_jobs, d_g, d_qoff, d_qsk, kmer, min_aa_len,
				d_rc, d_cnt, (const int64_t*)nullptr, (uint64_t*)nullptr);
		RFOK(hipGetLastError());
		RFOK(hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
		RFOK(hipStreamSynchronize(s));
		for (int32_t j = 0; j < n; ++j) woff[j + 1] = woff[j] + cnt[j];
		const int64_t n_w = woff[n];
		RFOK(hipMemcpyAsync(d_woff, woff.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMalloc((void**)&d_w, (size_t)(n_w + 1) * 8));
		RFOK(hipMalloc((void**)&d_ws, (size_t)(n_w + 1) * 8));
		hipLaunchKernelGGL(window_hits_kernel, dim3(n), dim3(256), 0, s, d_jobs, d_g, d_qoff, d_qsk, kmer, min_aa_len,
				d_rc, d_cnt, (const int64_t*)d_woff, d_w);
		RFOK(hipGetLastError());
		if (!sort_segments(d_w, d_ws, n_w, d_woff)) { nsh_set_error("refine: sorting the window k-mers failed"); goto done; }
		// the bucket pairs: count, offsets, write, sort per job
		hipLaunchKernelGGL(pairs_kernel, dim3(n), dim3(256), 0, s, d_jobs, d_woff, d_ws, d_qoff, d_qsk, max_ava, d_cnt,
				(const int64_t*)nullptr, (uint64_t*)nullptr);
		RFOK(hipGetLastError());
		RFOK(hipMemcpyAsync(pc.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, s));
		RFOK(hipStreamSynchronize(s));
		for (int32_t j = 0; j < n; ++j) p_off[j + 1] = p_off[j] + pc[j];
		const int64_t n_p = p_off[n];
		RFOK(hipMemcpyAsync(d_poff, p_off.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, s));
		RFOK(hipMalloc((void**)&d_p, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_ps, (size_t)(n_p + 1) * 8));
		hipLaunchKernelGGL(pairs_kernel, dim3(n), dim3(256), 0, s, d_jobs, d_woff, d_ws, d_qoff, d_qsk, max_ava, d_cnt,
				(const int64_t*)d_poff, d_p);
		RFOK(hipGetLastError());
		if (!sort_segments(d_p, d_ps, n_p, d_poff)) { nsh_set_error("refine: sorting the anchor pairs failed"); goto done; }
		nsh_count_launch(4);
		// mp_chain of every job's pairs (map.c:203), then the chains and their anchors to the host
		RFOK(hipMalloc((void**)&d_u, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_ao, (size_t)(n_p + 1) * 8));
		RFOK(hipMalloc((void**)&d_nu, (size_t)n * 4));
		if (n_p > 0) {
			if (!nsh_chain::run(s, n, d_poff, nullptr, pc.data(), d_ps, n_p, cp, d_nu, d_u, d_ao)) goto done;
		} else RFOK(hipMemsetAsync(d_nu, 0, (size_t)n * 4, s));
		u.resize((size_t)n_p), ao.resize((size_t)n_p);
		RFOK(hipMemcpyAsync(n_u.data(), d_nu, (size_t)n * 4, hipMemcpyDeviceToHost, s));
		if (n_p > 0) {
			if (!nsh_copy_to_host(u.data(), d_u, (size_t)n_p * 8, s)) goto done;
			if (!nsh_copy_to_host(ao.data(), d_ao, (size_t)n_p * 8, s)) goto done;
		}
		RFOK(hipStreamSynchronize(s));
		ok = 1;
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_jobs, d_qoff, d_cnt, d_woff, d_poff, d_qsk, d_w, d_ws, d_p, d_ps, d_u, d_ao, d_nu, d_tmp, d_rc };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	if (s) (void)hipStreamDestroy(s);
	return ok;
}
