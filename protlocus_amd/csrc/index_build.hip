// The k-mer block index of a genome (mp_idx_build without the file read, index.c:14-29, 52-97)
// on the device: block offsets per contig strand, the sketch of each strand (mp_sketch_nt4,
// sketch.c:108-169) and the bucketed entries (build_bidx).
//
// The reference sketches each strand with one host thread: it walks the three frames for clean
// ORFs (runs of at least min_aa_len stop- and N-free codons, the walk prescreen.c's ORF scan shares),
// hashes every k-mer of each run in the 13-letter alphabet, keeps those whose low mod_bit hash
// bits are zero, and sorts and deduplicates the (hash >> mod_bit) << 32 | block entries of the
// strand. The buckets are then filled strand by strand in that order.
//
// Here the whole genome goes to the device once; a reverse strand is complemented there. The runs
// come from the ORF walk kernels (orf_scan.hip, unsorted). One wave per run emits its k-mers, in
// a counting pass, a scan over the runs and a writing pass (no shared counter), every strand into
// one buffer. Block numbers are disjoint between strands and increase with the strand index, so
// one global radix sort and one unique pass give every bucket's entries in the reference's order,
// and a histogram and a scan give the bucket starts.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

// mp_hash32_mask (sketch.c:7-16): an invertible integer hash inside mask
__device__ __forceinline__ uint32_t hash32_mask(uint32_t key, uint32_t mask)
{
	key = (key + ~(key << 15)) & mask;
	key ^= key >> 10;
	key = (key + (key << 3)) & mask;
	key ^= key >> 6;
	key = (key + ~(key << 11)) & mask;
	key ^= key >> 16;
	return key;
}

__global__ void revcomp_kernel(const uint8_t *src, int64_t len, uint8_t *dst)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= len) return;
	const uint8_t c = src[len - 1 - k];
	dst[k] = c >= 4? c : (uint8_t)(3 - c);
}

struct SkArgs {
	const uint8_t *nt, *codon13;
	const int64_t *st; const int32_t *len; int64_t n_run;
	int32_t kmer, mod_bit, bbit;
	uint32_t mask_k, mask_mod;
	int64_t boff;
	int64_t *cnt;                 // per run: kept k-mers (counting pass, PASS 0)
	const int64_t *off;           // per run: first output slot (writing pass, PASS 1)
	uint64_t *out;
};

// one wave per run (mp_sketch_clean_orf, sketch.c:108-128), 64 codons per step: each lane
// translates one codon to its 13-letter code, and the k-mer closing at codon t (t >= kmer - 1),
// codons t-kmer+1..t at 4 bits each, the newest lowest, gathers the older codes from the lanes
// below by DPP shifts (lane 0 takes the previous step's lane 63). PASS 0 counts the kept k-mers
// of the run, PASS 1 writes them at the run's scanned offset.
template <int PASS>
__global__ void __launch_bounds__(256) sketch_kernel(SkArgs a)
{
	__shared__ uint8_t c13[64];
	if (threadIdx.x < 64) c13[threadIdx.x] = a.codon13[threadIdx.x];
	__syncthreads();
	const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
	if (r >= a.n_run) return;   // whole waves; no barrier follows
	const int L = threadIdx.x & 63;
	const int64_t st = a.st[r];
	const int32_t k = a.len[r];
	uint32_t last[6] = { 0, 0, 0, 0, 0, 0 };   // lane 63 of the previous step, per shift
	int64_t n = 0, o = PASS? a.off[r] : 0;
	for (int32_t t0 = 0; t0 < k; t0 += 64) {
		const int32_t t = t0 + L;
		uint32_t code = 0;
		if (t < k) {
			const uint8_t *c = a.nt + st + 3 * (int64_t)t;
			code = c13[c[0] << 4 | c[1] << 2 | c[2]];
		}
		uint32_t x = code, cur = code;
#pragma unroll
		for (int s = 1; s < 7; ++s) {
			if (s >= a.kmer) break;
			const uint32_t nx = (uint32_t)__builtin_amdgcn_update_dpp((int)last[s - 1], (int)cur, 0x138, 0xf, 0xf, false);
			last[s - 1] = (uint32_t)__builtin_amdgcn_readlane((int)cur, 63);
			cur = nx;
			x |= cur << (4 * s);
		}
		x &= a.mask_k;
		bool emit = false;
		uint32_t y = 0;
		if (t < k && t >= a.kmer - 1) {
			y = hash32_mask(x, a.mask_k);
			emit = (y & a.mask_mod) == 0;
		}
		const uint64_t bal = __ballot(emit);
		if (PASS == 0) {
			n += __popcll(bal);
		} else if (emit) {
			a.out[o + __popcll(bal & ((1ull << L) - 1))] =
				(uint64_t)(y >> a.mod_bit) << 32 | (uint32_t)(((st + 3 * (int64_t)t + 2) >> a.bbit) + a.boff);
		}
		if (PASS) o += __popcll(bal);
	}
	if (PASS == 0 && L == 0) a.cnt[r] = n;
}

__global__ void zero_u64_kernel(unsigned long long *p, int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) p[k] = 0;
}

// bucket sizes of the unique entries, and the block numbers
__global__ void bucket_kernel(const uint64_t *key, const int64_t *n_key, unsigned long long *cnt, uint32_t *kb)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= *n_key) return;
	const uint64_t x = key[k];
	atomicAdd(&cnt[x >> 32], 1ull);
	kb[k] = (uint32_t)x;
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

} // namespace

struct nsh_index_s {
	int32_t n_ctg = 0, kmer = 0, mod_bit = 0, bbit = 0, min_aa_len = 0;
	int64_t n_kb = 0, n_bucket = 0;
	uint32_t n_block = 0;
	std::vector<uint32_t> bo;
	int64_t *d_ki = nullptr;
	uint32_t *d_kb = nullptr;
};

#define IXOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("index: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" void nsh_index_destroy(nsh_index_t *ix)
{
	if (!ix) return;
	if (ix->d_ki) (void)hipFree(ix->d_ki);
	if (ix->d_kb) (void)hipFree(ix->d_kb);
	delete ix;
}

extern "C" nsh_index_t *nsh_index_build(int32_t n_ctg, const uint8_t *nt, const int64_t *ctg_off, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
{
	if (n_ctg < 0 || (n_ctg > 0 && (!nt || !ctg_off || !ctg_len)) || min_aa_len < 1 || kmer < 1 || kmer > 7 ||
			mod_bit < 0 || mod_bit >= 4 * kmer || bbit < 0 || bbit > 30) {
		nsh_set_error("index: bad arguments (need 1 <= kmer <= 7, 0 <= mod_bit < 4 kmer, 0 <= bbit <= 30, min_aa_len >= 1)");
		return nullptr;
	}
	if (!nsh_ensure_device()) return nullptr;
	nsh_index_t *ix = new nsh_index_s();
	ix->n_ctg = n_ctg, ix->kmer = kmer, ix->mod_bit = mod_bit, ix->bbit = bbit, ix->min_aa_len = min_aa_len;
	ix->n_bucket = (int64_t)1 << (kmer * 4 - mod_bit);
	// block offsets (mp_idx_boff, index.c:14-29)
	int64_t boff = 0, tot = 0, max_len = 0;
	ix->bo.resize((size_t)n_ctg * 2 + 1);
	for (int32_t c = 0; c < n_ctg; ++c) {
		if (ctg_off[c] < 0 || ctg_len[c] < 0) { nsh_set_error("index: contig %d out of range", c); delete ix; return nullptr; }
		const int64_t nb = (ctg_len[c] + ((int64_t)1 << bbit) - 1) >> bbit;
		ix->bo[2 * c] = (uint32_t)boff, boff += nb;
		ix->bo[2 * c + 1] = (uint32_t)boff, boff += nb;
		tot = std::max(tot, ctg_off[c] + ctg_len[c]), max_len = std::max(max_len, ctg_len[c]);
	}
	if (boff >= (int64_t)UINT32_MAX) { nsh_set_error("index: %lld blocks exceed 32 bits", (long long)boff); delete ix; return nullptr; }
	ix->bo[(size_t)n_ctg * 2] = ix->n_block = (uint32_t)boff;
	hipStream_t s = nullptr;
	uint8_t *d_g = nullptr, *d_rc = nullptr, *d_cod = nullptr, *d_cod13 = nullptr;
	uint64_t *d_raw = nullptr, *d_sorted = nullptr;
	unsigned long long *d_cnt = nullptr;
	int64_t *d_nuniq = nullptr;
	void *d_tmp = nullptr;
	nsh_orfs::Work ow;
	int64_t cap = 0, n_raw = 0, run_cap = 0, *d_rcnt = nullptr, *d_roff = nullptr;
	void *d_scan = nullptr;
	size_t scan_bytes = 0;
	for (int32_t c = 0; c < n_ctg; ++c) cap += 2 * (ctg_len[c] + 1);   // at most one k-mer per codon end and strand
	{
		IXOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		IXOK(hipMalloc((void**)&d_g, (size_t)tot + 8));
		IXOK(hipMalloc((void**)&d_rc, (size_t)max_len + 8));
		IXOK(hipMalloc((void**)&d_cod, 64));
		IXOK(hipMalloc((void**)&d_cod13, 64));
		IXOK(hipMalloc((void**)&d_raw, (size_t)cap * 8 + 8));
		if (tot) IXOK(hipMemcpyAsync(d_g, nt, (size_t)tot, hipMemcpyHostToDevice, s));
		IXOK(hipMemcpyAsync(d_cod, ns_tab_codon, 64, hipMemcpyHostToDevice, s));
		IXOK(hipMemcpyAsync(d_cod13, ns_tab_codon13, 64, hipMemcpyHostToDevice, s));
		for (int32_t j = 0; j < 2 * n_ctg; ++j) {   // contig strands (mp_ntseq_get orientation)
			const int64_t len = ctg_len[j >> 1];
			const uint8_t *seq = d_g + ctg_off[j >> 1];
			if (len < 3) continue;
			if (j & 1) {
				hipLaunchKernelGGL(revcomp_kernel, dim3(blocks(len)), dim3(256), 0, s, seq, len, d_rc);
				seq = d_rc;
			}
			int64_t n_run = 0;
			if (!nsh_orfs::find(s, seq, len, d_cod, min_aa_len, ow, &n_run)) goto fail;
			if (n_run == 0) continue;
			SkArgs a;
			a.nt = seq, a.codon13 = d_cod13, a.st = ow.st, a.len = ow.len, a.n_run = n_run;
			a.kmer = kmer, a.mod_bit = mod_bit, a.bbit = bbit;
			a.mask_k = (uint32_t)(((uint64_t)1 << (4 * kmer)) - 1), a.mask_mod = (1u << mod_bit) - 1;
			a.boff = ix->bo[j];
			// kept k-mers per run, their offsets after what the earlier strands wrote, then the k-mers
			if (n_run + 1 > run_cap) {
				IXOK(hipStreamSynchronize(s));
				if (d_rcnt) (void)hipFree(d_rcnt);
				if (d_roff) (void)hipFree(d_roff);
				d_rcnt = d_roff = nullptr;
				run_cap = n_run + 1 + n_run / 2;
				IXOK(hipMalloc((void**)&d_rcnt, (size_t)run_cap * 8));
				IXOK(hipMalloc((void**)&d_roff, (size_t)run_cap * 8));
			}
			a.cnt = d_rcnt, a.off = d_roff, a.out = d_raw + n_raw;
			hipLaunchKernelGGL(sketch_kernel<0>, dim3((unsigned)((n_run + 3) / 4)), dim3(256), 0, s, a);
			hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)(d_rcnt + n_run), (int64_t)1);
			size_t sb = 0;
			IXOK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, d_rcnt, d_roff, (int)(n_run + 1), s));
			if (sb > scan_bytes) {
				IXOK(hipStreamSynchronize(s));
				if (d_scan) (void)hipFree(d_scan);
				d_scan = nullptr;
				scan_bytes = sb + sb / 2;
				IXOK(hipMalloc(&d_scan, scan_bytes));
			}
			sb = scan_bytes;
			IXOK(hipcub::DeviceScan::ExclusiveSum(d_scan, sb, d_rcnt, d_roff, (int)(n_run + 1), s));
			int64_t n_strand = 0;
			IXOK(hipMemcpyAsync(&n_strand, d_roff + n_run, 8, hipMemcpyDeviceToHost, s));
			IXOK(hipStreamSynchronize(s));
			if (n_raw + n_strand > cap) { nsh_set_error("index: entries exceed the bound %lld", (long long)cap); goto fail; }
			hipLaunchKernelGGL(sketch_kernel<1>, dim3((unsigned)((n_run + 3) / 4)), dim3(256), 0, s, a);
			IXOK(hipGetLastError());
			n_raw += n_strand;
			nsh_count_launch(4 + (j & 1));
		}
		IXOK(hipStreamSynchronize(s));
		if (n_raw > INT_MAX) { nsh_set_error("index: %lld entries exceed one sort", (long long)n_raw); goto fail; }
		// one sort over every strand, unique entries, bucket sizes, bucket starts
		const int end_bit = 32 + 4 * kmer - mod_bit;
		size_t t1 = 0, t2 = 0, t3 = 0;
		IXOK(hipMalloc((void**)&d_sorted, (size_t)n_raw * 8 + 8));
		IXOK(hipMalloc((void**)&d_nuniq, 8));
		IXOK(hipMalloc((void**)&d_cnt, (size_t)(ix->n_bucket + 1) * 8));
		IXOK(hipMalloc((void**)&ix->d_ki, (size_t)(ix->n_bucket + 1) * 8));
		IXOK(hipcub::DeviceRadixSort::SortKeys(nullptr, t1, d_raw, d_sorted, (int)n_raw, 0, end_bit, s));
		IXOK(hipcub::DeviceSelect::Unique(nullptr, t2, d_sorted, d_raw, d_nuniq, (int)n_raw, s));
		IXOK(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, (const int64_t*)d_cnt, ix->d_ki, (int)(ix->n_bucket + 1), s));
		IXOK(hipMalloc(&d_tmp, std::max(t1, std::max(t2, t3))));
		size_t tb = std::max(t1, std::max(t2, t3));
		if (n_raw > 0) {
			IXOK(hipcub::DeviceRadixSort::SortKeys(d_tmp, tb, d_raw, d_sorted, (int)n_raw, 0, end_bit, s));
			tb = std::max(t1, std::max(t2, t3));
			IXOK(hipcub::DeviceSelect::Unique(d_tmp, tb, d_sorted, d_raw, d_nuniq, (int)n_raw, s));
		} else {
			hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)d_nuniq, (int64_t)1);
		}
		hipLaunchKernelGGL(zero_u64_kernel, dim3(blocks(ix->n_bucket + 1)), dim3(256), 0, s, d_cnt, ix->n_bucket + 1);
		IXOK(hipMemcpyAsync(&ix->n_kb, d_nuniq, 8, hipMemcpyDeviceToHost, s));
		IXOK(hipStreamSynchronize(s));
		IXOK(hipMalloc((void**)&ix->d_kb, (size_t)ix->n_kb * 4 + 8));
		if (ix->n_kb > 0)
			hipLaunchKernelGGL(bucket_kernel, dim3(blocks(ix->n_kb)), dim3(256), 0, s, d_raw, d_nuniq, d_cnt, ix->d_kb);
		tb = std::max(t1, std::max(t2, t3));
		IXOK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, (const int64_t*)d_cnt, ix->d_ki, (int)(ix->n_bucket + 1), s));
		IXOK(hipGetLastError());
		IXOK(hipStreamSynchronize(s));
		nsh_count_launch(6);
	}
	{
		void *bufs[] = { d_g, d_rc, d_cod, d_cod13, d_raw, d_sorted, d_cnt, d_nuniq, d_tmp, d_rcnt, d_roff,
			d_scan };
		for (void *b : bufs) if (b) (void)hipFree(b);
		ow.release();
		(void)hipStreamDestroy(s);
	}
	return ix;
fail:
	{
		if (s) (void)hipStreamSynchronize(s);
		void *bufs[] = { d_g, d_rc, d_cod, d_cod13, d_raw, d_sorted, d_cnt, d_nuniq, d_tmp, d_rcnt, d_roff,
			d_scan };
		for (void *b : bufs) if (b) (void)hipFree(b);
		ow.release();
		if (s) (void)hipStreamDestroy(s);
	}
	nsh_index_destroy(ix);
	return nullptr;
}

extern "C" int64_t nsh_index_size(const nsh_index_t *ix, int32_t what)
{
	if (!ix) return -1;
	return what == 0? ix->n_kb : what == 1? ix->n_bucket : what == 2? (int64_t)ix->n_block : -1;
}

extern "C" int nsh_index_copy(const nsh_index_t *ix, int64_t *ki, uint32_t *kb, uint32_t *bo)
{
	if (!ix) { nsh_set_error("index: null index"); return 0; }
	if (ki && hipMemcpy(ki, ix->d_ki, (size_t)ix->n_bucket * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		nsh_set_error("index: copy of the bucket starts failed");
		return 0;
	}
	if (kb && ix->n_kb > 0 && hipMemcpy(kb, ix->d_kb, (size_t)ix->n_kb * 4, hipMemcpyDeviceToHost) != hipSuccess) {
		nsh_set_error("index: copy of the entries failed");
		return 0;
	}
	if (bo) std::copy(ix->bo.begin(), ix->bo.end(), bo);
	return 1;
}
