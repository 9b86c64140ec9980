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

// ---- anchors of a batch of queries (map.c:583-597, default options) ----

struct AnArgs {
	const uint8_t *seq; const int64_t *qoff; const int32_t *qlen; int32_t n_q;
	const int64_t *sbase;       // per query: first slot of its seed region (q_len slots)
	const uint8_t *aa13;
	int32_t kmer, mod_bit;
	uint32_t mask_k, mask_mod;
	const int64_t *ki; const uint32_t *kb; int64_t n_bucket;
	uint64_t *seed;             // per query, at sbase: hash >> mod_bit << 32 | position, compacted
	uint64_t *cnt;              // per seed: its bucket size (then sorted per query)
	int32_t *n_seed, *max_occ;
	int32_t occ_cap;
	int64_t *n_anchor;          // per query: anchors kept
	const int64_t *aoff;        // per query: first anchor
	uint64_t *anchor;
};

// one wave per query: the protein sketch (mp_sketch_prot, sketch.c:18-38). Position i closes a
// k-mer when its last kmer residues are all in the 13-letter alphabet; the kept ones (low
// mod_bit hash bits zero) are compacted in order, with their bucket sizes.
__global__ void __launch_bounds__(64) qsketch_kernel(AnArgs a)
{
	const int q = blockIdx.x, L = threadIdx.x;
	const uint8_t *s = a.seq + a.qoff[q];
	const int32_t n = a.qlen[q];
	int32_t o = 0;
	for (int32_t i0 = 0; i0 < n; i0 += 64) {
		const int32_t i = i0 + L;
		bool emit = false;
		uint32_t y = 0;
		if (i < n && i >= a.kmer - 1) {
			uint32_t x = 0;
			bool ok = true;
			for (int t = 0; t < a.kmer; ++t) {
				const uint32_t c = a.aa13[s[i - a.kmer + 1 + t]];
				ok = ok && c < 14;
				x = (x << 4 | c) & a.mask_k;
			}
			if (ok) {
				y = hash32_mask(x, a.mask_k);
				emit = (y & a.mask_mod) == 0;
			}
		}
		const uint64_t bal = __ballot(emit);
		if (emit) {
			const int64_t k = a.sbase[q] + o + __popcll(bal & ((1ull << L) - 1));
			const uint64_t b = y >> a.mod_bit;
			a.seed[k] = b << 32 | (uint32_t)i;
			a.cnt[k] = (uint64_t)(a.ki[b + 1] - a.ki[b]);
		}
		o += __popcll(bal);
	}
	if (L == 0) a.n_seed[q] = o;
}

// per query: the repeat cutoff (mp_cal_max_occ, map.c:304-318) from its sorted bucket sizes,
// capped at the option's max_occ; fewer than 8 seeds keep the option
__global__ void occ_kernel(AnArgs a)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= a.n_q) return;
	const int32_t n = a.n_seed[q];
	int32_t occ = a.occ_cap;
	if (n >= 8) {
		const uint64_t *c = a.cnt + a.sbase[q];
		const uint64_t q25 = c[(int64_t)(n * .25 + .499)], q75 = c[(int64_t)(n * .75 + .499)];
		const int32_t v = (int32_t)(q75 + (q75 - q25) * 1.5 + 10.);
		occ = v < occ? v : occ;
	}
	a.max_occ[q] = occ;
}

// one wave per query: bucket entries of the seeds under the cutoff (mp_collect_anchors,
// map.c:336-395, no repeat rescue); PASS 0 counts them, PASS 1 writes
// kb << 32 | position at the query's offset
template <int PASS>
__global__ void __launch_bounds__(64) collect_kernel(AnArgs a)
{
	const int q = blockIdx.x, L = threadIdx.x;
	const int32_t n = a.n_seed[q], occ = a.max_occ[q];
	const uint64_t *sd = a.seed + a.sbase[q];
	int64_t o = PASS? a.aoff[q] : 0;
	for (int32_t i0 = 0; i0 < n; i0 += 64) {
		const int32_t i = i0 + L;
		int64_t st = 0, take = 0;
		if (i < n) {
			const uint64_t b = sd[i] >> 32;
			st = a.ki[b];
			take = a.ki[b + 1] - st;
			if (take > occ) take = 0;
		}
		// inclusive scan of the takes across the wave
		int64_t inc = take;
		for (int d = 1; d < 64; d <<= 1) {
			const int64_t t = __shfl_up(inc, d);
			if (L >= d) inc += t;
		}
		if (PASS) {
			const uint32_t pos = i < n? (uint32_t)sd[i] : 0;
			for (int64_t j = 0; j < take; ++j)
				a.anchor[o + inc - take + j] = (uint64_t)a.kb[st + j] << 32 | pos;
		}
		o += __shfl(inc, 63);
	}
	if (!PASS && L == 0) a.n_anchor[q] = o;
}

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

namespace {
// the reference's 4-bit genome (mp_ntdb_t.seq: base i in the low nibble of byte i/2 when i is
// even, the high nibble when odd) to one code per byte
__global__ void unpack4_kernel(const uint8_t *pk, int64_t n, uint8_t *out)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (2 * i >= n) return;
	const uint8_t b = pk[i];
	out[2 * i] = b & 0xf;
	if (2 * i + 1 < n) out[2 * i + 1] = b >> 4;
}

nsh_index_t *index_build(int32_t n_ctg, const uint8_t *nt, bool packed, const int64_t *ctg_off,
		const int64_t *ctg_len, int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit);
} // namespace

extern "C" nsh_index_t *nsh_index_build(int32_t n_ctg, const uint8_t *nt, const int64_t *ctg_off, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
{
	return index_build(n_ctg, nt, false, ctg_off, ctg_len, min_aa_len, kmer, mod_bit, bbit);
}

extern "C" nsh_index_t *nsh_index_build_nt4(int32_t n_ctg, const uint8_t *seq4, const int64_t *ctg_off,
		const int64_t *ctg_len, int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
{
	return index_build(n_ctg, seq4, true, ctg_off, ctg_len, min_aa_len, kmer, mod_bit, bbit);
}

namespace {
nsh_index_t *index_build(int32_t n_ctg, const uint8_t *nt, bool packed, const int64_t *ctg_off,
		const int64_t *ctg_len, int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
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
	uint8_t *d_g = nullptr, *d_rc = nullptr, *d_cod = nullptr, *d_cod13 = nullptr, *d_pk = nullptr;
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
		if (tot && !packed && !nsh_copy_to_device(d_g, nt, (size_t)tot, s)) goto fail;
		if (tot && packed) {   // half the bytes cross PCIe; the device unpacks them
			IXOK(hipMalloc((void**)&d_pk, (size_t)(tot + 1) / 2 + 8));
			if (!nsh_copy_to_device(d_pk, nt, (size_t)(tot + 1) / 2, s)) goto fail;
			hipLaunchKernelGGL(unpack4_kernel, dim3((unsigned)(((tot + 1) / 2 + 255) / 256)), dim3(256), 0, s, d_pk, tot, d_g);
			IXOK(hipGetLastError());
		}
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
			d_scan, d_pk };
		for (void *b : bufs) if (b) (void)hipFree(b);
		ow.release();
		(void)hipStreamDestroy(s);
	}
	return ix;
fail:
	{
		if (s) (void)hipStreamSynchronize(s);
		void *bufs[] = { d_g, d_rc, d_cod, d_cod13, d_raw, d_sorted, d_cnt, d_nuniq, d_tmp, d_rcnt, d_roff,
			d_scan, d_pk };
		for (void *b : bufs) if (b) (void)hipFree(b);
		ow.release();
		if (s) (void)hipStreamDestroy(s);
	}
	nsh_index_destroy(ix);
	return nullptr;
}
} // namespace

extern "C" int64_t nsh_index_size(const nsh_index_t *ix, int32_t what)
{
	if (!ix) return -1;
	return what == 0? ix->n_kb : what == 1? ix->n_bucket : what == 2? (int64_t)ix->n_block : -1;
}

extern "C" int nsh_index_copy(const nsh_index_t *ix, int64_t *ki, uint32_t *kb, uint32_t *bo)
{
	if (!ix) { nsh_set_error("index: null index"); return 0; }
	if (ki && !nsh_copy_to_host(ki, ix->d_ki, (size_t)ix->n_bucket * 8, nullptr)) {
		nsh_set_error("index: copy of the bucket starts failed");
		return 0;
	}
	if (kb && ix->n_kb > 0 && !nsh_copy_to_host(kb, ix->d_kb, (size_t)ix->n_kb * 4, nullptr)) {
		nsh_set_error("index: copy of the entries failed");
		return 0;
	}
	if (bo) std::copy(ix->bo.begin(), ix->bo.end(), bo);
	return 1;
}

#define ANOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("anchors: %s: %s", #x, hipGetErrorString(e_)); goto done; } } while (0)

namespace {
int check_queries(int32_t n_q, const int64_t *q_off, const int32_t *q_len, int64_t *l_seq, std::vector<int64_t> &sbase)
{
	*l_seq = 0;
	sbase.assign(n_q + 1, 0);
	for (int32_t q = 0; q < n_q; ++q) {
		if (q_off[q] < 0 || q_len[q] < 0) { nsh_set_error("anchors: query %d out of range", q); return 0; }
		*l_seq = std::max(*l_seq, q_off[q] + q_len[q]);
		sbase[q + 1] = sbase[q] + q_len[q];
	}
	if (sbase[n_q] > INT_MAX) {
		nsh_set_error("anchors: %lld query residues exceed one sort", (long long)sbase[n_q]);
		return 0;
	}
	return 1;
}

// every query's anchors (map.c:583-597), sorted, left on the device: *pd_anch (aoff[n_q] of them,
// query q's at aoff[q]) and *pd_aoff (aoff on the device), both for the caller to free. Waits for
// stream s; 1 on success.
int collect_anchors(const nsh_index_t *ix, int32_t n_q, const char *seq, const int64_t *q_off, const int32_t *q_len,
		int32_t max_occ, hipStream_t s, std::vector<int64_t> &aoff, int64_t **pd_aoff, uint64_t **pd_anch)
{
	int64_t l_seq = 0, l_seed = 0;
	std::vector<int64_t> sbase;
	if (!check_queries(n_q, q_off, q_len, &l_seq, sbase)) return 0;
	l_seed = sbase[n_q];
	int ok = 0;
	uint8_t *d_seq = nullptr, *d_aa13 = nullptr;
	int64_t *d_qoff = nullptr, *d_nanch = nullptr, *d_aoff = nullptr, *d_soff = nullptr, *d_send = nullptr,
		*d_sbase = nullptr;
	int32_t *d_qlen = nullptr, *d_nseed = nullptr, *d_occ = nullptr;
	uint64_t *d_seed = nullptr, *d_cnt = nullptr, *d_cnt2 = nullptr, *d_anch = nullptr, *d_anch2 = nullptr;
	void *d_tmp = nullptr;
	size_t tb = 0, t2 = 0;
	std::vector<int64_t> nan(n_q), soff(n_q), send(n_q);
	std::vector<int32_t> nseed(n_q);
	AnArgs a = {};
	aoff.assign(n_q + 1, 0);
	nsh_step_clock clk("collect_anchors");
	{
		ANOK(hipMalloc((void**)&d_seq, (size_t)l_seq + 8));
		ANOK(hipMalloc((void**)&d_aa13, 256));
		ANOK(hipMalloc((void**)&d_qoff, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_qlen, (size_t)n_q * 4));
		ANOK(hipMalloc((void**)&d_nseed, (size_t)n_q * 4));
		ANOK(hipMalloc((void**)&d_occ, (size_t)n_q * 4));
		ANOK(hipMalloc((void**)&d_nanch, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_aoff, (size_t)(n_q + 1) * 8));
		ANOK(hipMalloc((void**)&d_soff, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_send, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_sbase, (size_t)(n_q + 1) * 8));
		ANOK(hipMalloc((void**)&d_seed, (size_t)l_seed * 8 + 8));
		ANOK(hipMalloc((void**)&d_cnt, (size_t)l_seed * 8 + 8));
		ANOK(hipMalloc((void**)&d_cnt2, (size_t)l_seed * 8 + 8));
		ANOK(hipMemcpyAsync(d_sbase, sbase.data(), (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		if (l_seq) ANOK(hipMemcpyAsync(d_seq, seq, (size_t)l_seq, hipMemcpyHostToDevice, s));
		ANOK(hipMemcpyAsync(d_aa13, ns_tab_aa13, 256, hipMemcpyHostToDevice, s));
		ANOK(hipMemcpyAsync(d_qoff, q_off, (size_t)n_q * 8, hipMemcpyHostToDevice, s));
		ANOK(hipMemcpyAsync(d_qlen, q_len, (size_t)n_q * 4, hipMemcpyHostToDevice, s));
		a.seq = d_seq, a.qoff = d_qoff, a.qlen = d_qlen, a.n_q = n_q, a.sbase = d_sbase, a.aa13 = d_aa13;
		a.kmer = ix->kmer, a.mod_bit = ix->mod_bit;
		a.mask_k = (uint32_t)(((uint64_t)1 << (4 * ix->kmer)) - 1), a.mask_mod = (1u << ix->mod_bit) - 1;
		a.ki = ix->d_ki, a.kb = ix->d_kb, a.n_bucket = ix->n_bucket;
		a.seed = d_seed, a.cnt = d_cnt, a.n_seed = d_nseed, a.max_occ = d_occ, a.occ_cap = max_occ;
		a.n_anchor = d_nanch, a.aoff = d_aoff;
		hipLaunchKernelGGL(qsketch_kernel, dim3(n_q), dim3(64), 0, s, a);
		ANOK(hipGetLastError());
		ANOK(hipMemcpyAsync(nseed.data(), d_nseed, (size_t)n_q * 4, hipMemcpyDeviceToHost, s));
		ANOK(hipStreamSynchronize(s));
		clk.lap("upload and seeds", l_seed);
		// the bucket sizes of each query's seeds, sorted (segments of the per-query regions)
		for (int32_t q = 0; q < n_q; ++q) soff[q] = sbase[q], send[q] = sbase[q] + nseed[q];
		ANOK(hipMemcpyAsync(d_soff, soff.data(), (size_t)n_q * 8, hipMemcpyHostToDevice, s));
		ANOK(hipMemcpyAsync(d_send, send.data(), (size_t)n_q * 8, hipMemcpyHostToDevice, s));
		ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, d_cnt, d_cnt2, (int)l_seed, n_q, d_soff, d_send,
				0, 64, s));
		ANOK(hipMalloc(&d_tmp, tb + 8));
		ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, tb, d_cnt, d_cnt2, (int)l_seed, n_q, d_soff, d_send,
				0, 64, s));
		a.cnt = d_cnt2;
		hipLaunchKernelGGL(occ_kernel, dim3((n_q + 255) / 256), dim3(256), 0, s, a);
		hipLaunchKernelGGL(collect_kernel<0>, dim3(n_q), dim3(64), 0, s, a);
		ANOK(hipGetLastError());
		ANOK(hipMemcpyAsync(nan.data(), d_nanch, (size_t)n_q * 8, hipMemcpyDeviceToHost, s));
		ANOK(hipStreamSynchronize(s));
		clk.lap("seed sort and counts");
		for (int32_t q = 0; q < n_q; ++q) aoff[q + 1] = aoff[q] + nan[q];
		const int64_t n_a = aoff[n_q];
		if (n_a > INT_MAX) { nsh_set_error("anchors: %lld anchors exceed one sort", (long long)n_a); goto done; }
		ANOK(hipMemcpyAsync(d_aoff, aoff.data(), (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		ANOK(hipMalloc((void**)&d_anch, (size_t)n_a * 8 + 8));
		ANOK(hipMalloc((void**)&d_anch2, (size_t)n_a * 8 + 8));
		a.anchor = d_anch;
		hipLaunchKernelGGL(collect_kernel<1>, dim3(n_q), dim3(64), 0, s, a);
		ANOK(hipGetLastError());
		// each query's anchors, sorted (radix_sort_mp64 in map.c:596)
		ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, t2, d_anch, d_anch2, (int)n_a, n_q, d_aoff,
				d_aoff + 1, 0, 64, s));
		if (t2 > tb) {
			ANOK(hipStreamSynchronize(s));
			(void)hipFree(d_tmp), d_tmp = nullptr;
			ANOK(hipMalloc(&d_tmp, t2));
			tb = t2;
		}
		ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, tb, d_anch, d_anch2, (int)n_a, n_q, d_aoff, d_aoff + 1,
				0, 64, s));
		ANOK(hipStreamSynchronize(s));
		nsh_count_launch(6);
		clk.lap("anchors and sort");
		*pd_aoff = d_aoff, *pd_anch = d_anch2;
		d_aoff = nullptr, d_anch2 = nullptr;
		ok = 1;
	}
done:
	(void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_seq, d_aa13, d_qoff, d_qlen, d_nseed, d_occ, d_nanch, d_aoff, d_soff, d_send, d_seed, d_cnt,
			d_cnt2, d_anch, d_anch2, d_tmp, d_sbase };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	clk.lap("free");
	return ok;
}
} // namespace

extern "C" int64_t nsh_index_anchors(const nsh_index_t *ix, int32_t n_q, const char *seq, const int64_t *q_off,
		const int32_t *q_len, int32_t max_occ, int64_t *a_off, uint64_t *anchors, int64_t max_anchors)
{
	if (!ix || n_q < 0 || (n_q > 0 && (!seq || !q_off || !q_len || !a_off)) || max_anchors < 0 ||
			(max_anchors > 0 && !anchors)) {
		nsh_set_error("anchors: bad arguments");
		return -1;
	}
	if (n_q == 0) { if (a_off) a_off[0] = 0; return 0; }
	int64_t total = -1;
	hipStream_t s = nullptr;
	int64_t *d_aoff = nullptr;
	uint64_t *d_anch = nullptr;
	std::vector<int64_t> aoff;
	nsh_step_clock clk("nsh_index_chain");
	{
		ANOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		if (!collect_anchors(ix, n_q, seq, q_off, q_len, max_occ, s, aoff, &d_aoff, &d_anch)) goto done;
		clk.lap("anchors", aoff[n_q]);
		const int64_t n_a = aoff[n_q];
		if (max_anchors > 0 && n_a > 0)
			if (!nsh_copy_to_host(anchors, d_anch, (size_t)std::min(n_a, max_anchors) * 8, s)) goto done;
		ANOK(hipStreamSynchronize(s));
		std::copy(aoff.begin(), aoff.end(), a_off);
		total = n_a;
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	if (d_aoff) (void)hipFree(d_aoff);
	if (d_anch) (void)hipFree(d_anch);
	if (s) (void)hipStreamDestroy(s);
	return total;
}


namespace {
// per query: the anchors held by the chains u[off[q], off[q] + nu[q]); end[q] = off[q] + that
__global__ void kept_kernel(int32_t n_q, const int64_t *off, const int32_t *nu, const uint64_t *u, int64_t *kept,
		int64_t *end)
{
	const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n_q) return;
	int64_t m = 0;
	for (int32_t i = 0; i < nu[q]; ++i) m += (uint32_t)u[off[q] + i];
	kept[q] = m;
	end[q] = off[q] + m;
}

// each query's chains and chain anchors from its region into the packed outputs
__global__ void pack_chains_kernel(const int64_t *off, const int32_t *nu, const int64_t *kept, const int64_t *uoff,
		const int64_t *coff, const uint64_t *u, const uint64_t *ao, uint64_t *uc, uint64_t *ac)
{
	const int32_t q = blockIdx.x;
	for (int64_t i = threadIdx.x; i < nu[q]; i += blockDim.x) uc[uoff[q] + i] = u[off[q] + i];
	for (int64_t i = threadIdx.x; i < kept[q]; i += blockDim.x) ac[coff[q] + i] = ao[off[q] + i];
}
} // namespace

extern "C" int64_t nsh_index_chain(const nsh_index_t *ix, int32_t n_q, const char *seq, const int64_t *q_off,
		const int32_t *q_len, int32_t max_occ, const nsh_chain_opt_t *opt, int64_t *u_off, uint64_t *u,
		int64_t max_u, int64_t *c_off, uint64_t *anchors, int64_t max_anchors)
{
	if (!ix || !opt || n_q < 0 || (n_q > 0 && (!seq || !q_off || !q_len)) || !u_off || !c_off || max_u < 0 ||
			max_anchors < 0 || (max_u > 0 && !u) || (max_anchors > 0 && !anchors)) {
		nsh_set_error("index chain: bad arguments");
		return -1;
	}
	u_off[0] = c_off[0] = 0;
	if (n_q == 0) return 0;
	int64_t total = -1;
	hipStream_t s = nullptr;
	int64_t *d_aoff = nullptr, *d_kept = nullptr, *d_end = nullptr, *d_uoff = nullptr, *d_coff = nullptr;
	uint64_t *d_anch = nullptr, *d_u = nullptr, *d_ao = nullptr, *d_uc = nullptr, *d_ac = nullptr;
	int32_t *d_nu = nullptr;
	void *d_tmp = nullptr;
	size_t tb = 0;
	std::vector<int64_t> aoff, cnt(n_q), kept(n_q);
	std::vector<int32_t> nu(n_q);
	const int32_t w = 1 << ix->bbit;
	const nsh_chain::Params pre = { w, w, w, opt->max_skip, opt->max_iter, 2, 0, opt->chn_coef_log, opt->is_spliced,
		ix->kmer, ix->bbit };
	const nsh_chain::Params main = { opt->max_dist_x, opt->max_dist_y, opt->bw, opt->max_skip, opt->max_iter,
		opt->min_cnt, opt->min_sc, opt->chn_coef_log, opt->is_spliced, ix->kmer, ix->bbit };
	nsh_step_clock clk("nsh_index_chain");
	{
		ANOK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		if (!collect_anchors(ix, n_q, seq, q_off, q_len, max_occ, s, aoff, &d_aoff, &d_anch)) goto done;
		clk.lap("anchors", aoff[n_q]);
		const int64_t n_a = aoff[n_q];
		const size_t na = (size_t)(n_a > 0? n_a : 1);
		for (int32_t q = 0; q < n_q; ++q) cnt[q] = aoff[q + 1] - aoff[q];
		ANOK(hipMalloc((void**)&d_nu, (size_t)n_q * 4));
		ANOK(hipMalloc((void**)&d_kept, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_end, (size_t)n_q * 8));
		ANOK(hipMalloc((void**)&d_u, na * 8));
		ANOK(hipMalloc((void**)&d_ao, na * 8));
		const int64_t *d_cnt = nullptr, *h_cnt = cnt.data();
		if (opt->pre_chain && opt->is_spliced) {   // map.c:611-623
			if (!nsh_chain::run(s, n_q, d_aoff, nullptr, cnt.data(), d_anch, n_a, pre, d_nu, d_u, d_ao)) goto done;
			hipLaunchKernelGGL(kept_kernel, dim3((n_q + 255) / 256), dim3(256), 0, s, n_q, d_aoff, d_nu, d_u, d_kept,
					d_end);
			ANOK(hipGetLastError());
			ANOK(hipMemcpyAsync(kept.data(), d_kept, (size_t)n_q * 8, hipMemcpyDeviceToHost, s));
			ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, d_ao, d_anch, (int)n_a, n_q, d_aoff, d_end,
					0, 64, s));
			ANOK(hipMalloc(&d_tmp, tb + 8));
			ANOK(hipcub::DeviceSegmentedRadixSort::SortKeys(d_tmp, tb, d_ao, d_anch, (int)n_a, n_q, d_aoff, d_end,
					0, 64, s));
			ANOK(hipStreamSynchronize(s));
			d_cnt = d_kept, h_cnt = kept.data();
			nsh_count_launch(3);
			clk.lap("pre-chain and sort");
		}
		if (!nsh_chain::run(s, n_q, d_aoff, d_cnt, h_cnt, d_anch, n_a, main, d_nu, d_u, d_ao)) goto done;
		hipLaunchKernelGGL(kept_kernel, dim3((n_q + 255) / 256), dim3(256), 0, s, n_q, d_aoff, d_nu, d_u, d_kept,
				d_end);
		ANOK(hipGetLastError());
		ANOK(hipMemcpyAsync(nu.data(), d_nu, (size_t)n_q * 4, hipMemcpyDeviceToHost, s));
		ANOK(hipMemcpyAsync(kept.data(), d_kept, (size_t)n_q * 8, hipMemcpyDeviceToHost, s));
		ANOK(hipStreamSynchronize(s));
		clk.lap("chain");
		for (int32_t q = 0; q < n_q; ++q) u_off[q + 1] = u_off[q] + nu[q], c_off[q + 1] = c_off[q] + kept[q];
		const int64_t n_uc = u_off[n_q], n_ac = c_off[n_q];
		ANOK(hipMalloc((void**)&d_uoff, (size_t)(n_q + 1) * 8));
		ANOK(hipMalloc((void**)&d_coff, (size_t)(n_q + 1) * 8));
		ANOK(hipMalloc((void**)&d_uc, (size_t)n_uc * 8 + 8));
		ANOK(hipMalloc((void**)&d_ac, (size_t)n_ac * 8 + 8));
		ANOK(hipMemcpyAsync(d_uoff, u_off, (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		ANOK(hipMemcpyAsync(d_coff, c_off, (size_t)(n_q + 1) * 8, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(pack_chains_kernel, dim3(n_q), dim3(256), 0, s, d_aoff, d_nu, d_kept, d_uoff, d_coff, d_u,
				d_ao, d_uc, d_ac);
		ANOK(hipGetLastError());
		if (max_u > 0 && n_uc > 0)
			ANOK(hipMemcpyAsync(u, d_uc, (size_t)std::min(n_uc, max_u) * 8, hipMemcpyDeviceToHost, s));
		if (max_anchors > 0 && n_ac > 0)
			ANOK(hipMemcpyAsync(anchors, d_ac, (size_t)std::min(n_ac, max_anchors) * 8, hipMemcpyDeviceToHost, s));
		ANOK(hipStreamSynchronize(s));
		nsh_count_launch(3);
		total = n_uc;
		clk.lap("pack and download", n_ac);
	}
done:
	if (s) (void)hipStreamSynchronize(s);
	{
		void *bufs[] = { d_aoff, d_kept, d_end, d_uoff, d_coff, d_anch, d_u, d_ao, d_uc, d_ac, d_nu, d_tmp };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	if (s) (void)hipStreamDestroy(s);
	clk.lap("free");
	return total;
}
