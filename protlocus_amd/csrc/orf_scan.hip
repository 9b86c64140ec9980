// Six-frame ORF walk of ps_orf_worker + ps_add_orf (prescreen.c:87-132) on the device.
//
// The reference walks the sequence once, keeping one open run per frame. A run grows by each
// valid non-stop codon. A stop codon closes its frame's run; an N closes all three. A closed run
// of at least min_orf codons is emitted, and at an N or at the end of the sequence the frames are
// emitted in order 0, 1, 2.
//
// Here the codons of each frame p (codon c ends at base 3c + (p + 2) % 3) are classified
// independently: good (valid, not a stop), stop, or invalid (a base that is not ACGT, or fewer
// than three bases). A run is a maximal stretch of good codons. Its start is the last non-good
// codon before it plus one: a per-chunk maximum, a scan across chunks, and a scan inside the
// chunk. Its emission position is computed locally:
//   * the end of the next codon when that codon is a stop;
//   * otherwise the first non-ACGT base after the run, which lies within the next three bases
//     or the sequence's last two;
//   * otherwise the sequence end.
// Sorting the emitted runs by (position, frame) reproduces the reference's order exactly. The
// residues of the sorted runs are written back-to-back, as ps_add_orf appends them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <chrono>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int NT = 256, PER = 16, CH = NT * PER;   // codons per chunk and frame

struct OrfArgs {
	const uint8_t *nt; int64_t len;
	const uint8_t *codon;
	int32_t min_orf;
	int64_t ncod[3], nchunk[3], chunk0[3];   // codons, chunks, first chunk (block) of each frame
	int64_t *lastbad;                        // per chunk: last non-good codon index, or -1
	int64_t *carry;                          // per chunk: last non-good codon index before it, or -1
	unsigned long long *count;               // emitted runs
	int64_t cap;
	uint64_t *key; uint32_t *idx;            // per run: position * 4 + frame, record index
	int64_t *st; int32_t *n_aa;              // per run: first base, codons
};

#define DEV static __device__ __forceinline__

// 0 good, 1 stop, 2 invalid; *aa the residue of a valid codon
DEV int status(const OrfArgs &a, int p, int64_t c)
{
	const int64_t i = 3 * c + (p + 2) % 3;
	if (i < 2 || i >= a.len) return 2;
	const int b0 = a.nt[i - 2], b1 = a.nt[i - 1], b2 = a.nt[i];
	if (b0 > 3 || b1 > 3 || b2 > 3) return 2;
	return a.codon[b0 << 4 | b1 << 2 | b2] >= 20? 1 : 0;
}

DEV void locate(const OrfArgs &a, int &p, int64_t &ch)
{
	const int64_t b = blockIdx.x;
	p = b < a.chunk0[1]? 0 : b < a.chunk0[2]? 1 : 2;
	ch = b - a.chunk0[p];
}

__global__ void __launch_bounds__(NT) orf_summary_kernel(OrfArgs a)
{
	int p;
	int64_t ch;
	locate(a, p, ch);
	const int64_t c0 = ch * CH;
	int64_t last = -1;
	for (int u = 0; u < PER; ++u) {
		const int64_t c = c0 + (int64_t)u * NT + threadIdx.x;
		if (c < a.ncod[p] && status(a, p, c) != 0) last = c > last? c : last;
	}
	typedef hipcub::BlockReduce<int64_t, NT> Red;
	__shared__ typename Red::TempStorage tmp;
	const int64_t m = Red(tmp).Reduce(last, hipcub::Max());
	if (threadIdx.x == 0) a.lastbad[blockIdx.x] = m;
}

// one block per frame: exclusive running maximum over the frame's chunks, a tile of NT chunks
// at a time with the maximum of the earlier tiles carried along
__global__ void __launch_bounds__(NT) orf_carry_kernel(OrfArgs a)
{
	const int p = blockIdx.x;
	typedef hipcub::BlockScan<int64_t, NT> Scan;
	__shared__ typename Scan::TempStorage tmp;
	int64_t run = -1;
	for (int64_t t0 = 0; t0 < a.nchunk[p]; t0 += NT) {
		const int64_t ch = t0 + threadIdx.x;
		const int64_t v = ch < a.nchunk[p]? a.lastbad[a.chunk0[p] + ch] : -1;
		int64_t before, tile;
		Scan(tmp).ExclusiveScan(v, before, (int64_t)-1, hipcub::Max(), tile);
		if (ch < a.nchunk[p]) a.carry[a.chunk0[p] + ch] = before > run? before : run;
		run = tile > run? tile : run;
		__syncthreads();   // the scan storage is reused by the next tile
	}
}

__global__ void __launch_bounds__(NT) orf_emit_kernel(OrfArgs a)
{
	int p;
	int64_t ch;
	locate(a, p, ch);
	const int64_t cb = ch * CH + (int64_t)threadIdx.x * PER;   // this thread's consecutive codons
	int8_t s[PER + 1];
	int64_t last = -1;
#pragma unroll
	for (int u = 0; u <= PER; ++u) {
		const int64_t c = cb + u;
		s[u] = c < a.ncod[p]? (int8_t)status(a, p, c) : (int8_t)2;
		if (u < PER && s[u] != 0 && c < a.ncod[p]) last = c;
	}
	typedef hipcub::BlockScan<int64_t, NT> Scan;
	typedef hipcub::BlockScan<int32_t, NT> Count;
	__shared__ typename Scan::TempStorage tmp;
	__shared__ typename Count::TempStorage tmp2;
	__shared__ unsigned long long base;
	int64_t before;
	Scan(tmp).ExclusiveScan(last, before, (int64_t)-1, hipcub::Max());
	const int64_t carry = a.carry[blockIdx.x];
	const int64_t bad0 = before > carry? before : carry;   // last non-good codon before cb
	// runs ending in this thread's codons: counted, then one atomic per block places them
	int32_t n_emit = 0;
	{
		int64_t bad = bad0;
#pragma unroll
		for (int u = 0; u < PER; ++u) {
			const int64_t c = cb + u;
			if (c >= a.ncod[p]) break;
			if (s[u] != 0) { bad = c; continue; }
			if (s[u + 1] == 0) continue;
			if (c - bad >= a.min_orf) ++n_emit;
		}
	}
	int32_t off, total;
	Count(tmp2).ExclusiveSum(n_emit, off, total);
	if (threadIdx.x == 0) base = total? atomicAdd(a.count, (unsigned long long)total) : 0ull;
	__syncthreads();
	unsigned long long slot = base + (unsigned long long)off;
	int64_t bad = bad0;
#pragma unroll
	for (int u = 0; u < PER; ++u) {
		const int64_t c = cb + u;
		if (c >= a.ncod[p]) break;
		if (s[u] != 0) { bad = c; continue; }
		if (s[u + 1] == 0) continue;                 // the run goes on
		const int64_t k = c - bad;                   // codons of the run ending at c
		if (k < a.min_orf) continue;
		const int64_t e = 3 * c + (p + 2) % 3;       // its last base
		int64_t ev;
		if (c + 1 < a.ncod[p] && s[u + 1] == 1) ev = e + 3;
		else {
			ev = a.len;
			const int64_t lim = e + 3 < a.len? e + 3 : a.len - 1;
			for (int64_t x = e + 1; x <= lim; ++x)
				if (a.nt[x] > 3) { ev = x; break; }
		}
		if ((int64_t)slot < a.cap) {
			a.key[slot] = (uint64_t)ev * 4 + (uint64_t)p;
			a.idx[slot] = (uint32_t)slot;
			a.st[slot] = e + 1 - 3 * k;
			a.n_aa[slot] = (int32_t)k;
		}
		++slot;
	}
}

// sorted runs: first base, codons and residues (one wave per run)
__global__ void __launch_bounds__(256) orf_gather_kernel(const OrfArgs a, const uint32_t *order, int64_t n,
		const int64_t *aa_off, int64_t *st_out, int32_t *len_out, uint8_t *aa)
{
	const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
	if (r >= n) return;
	const uint32_t k = order[r];
	const int64_t st = a.st[k];
	const int32_t n_aa = a.n_aa[k];
	if ((threadIdx.x & 63) == 0) st_out[r] = st, len_out[r] = n_aa;
	uint8_t *o = aa + aa_off[r];
	for (int t = threadIdx.x & 63; t < n_aa; t += 64) {
		const uint8_t *cdn = a.nt + st + 3 * (int64_t)t;
		o[t] = a.codon[cdn[0] << 4 | cdn[1] << 2 | cdn[2]];
	}
}

std::mutex g_mu;

__global__ void zero_count_kernel(unsigned long long *p)
{
	if (threadIdx.x == 0) *p = 0;
}

// the frame layout of a sequence of len bases: codons, chunks and first chunk of each frame
int64_t layout(OrfArgs &a, int64_t len, int32_t min_orf)
{
	int64_t nblk = 0;
	for (int p = 0; p < 3; ++p) {
		const int64_t r = (p + 2) % 3;
		a.ncod[p] = len > r? (len - r + 2) / 3 : 0;
		a.nchunk[p] = (a.ncod[p] + CH - 1) / CH;
		a.chunk0[p] = nblk;
		nblk += a.nchunk[p];
	}
	a.len = len, a.min_orf = min_orf;
	a.cap = 3 * (len / 3 / min_orf + 2);
	return nblk;
}

} // namespace

void nsh_orfs::Work::release()
{
	for (int i = 0; i < 6; ++i) if (p[i]) (void)hipFree(p[i]), p[i] = nullptr, n[i] = 0;
	st = nullptr, len = nullptr;
}

int nsh_orfs::find(hipStream_t s, const uint8_t *d_nt, int64_t len, const uint8_t *d_codon, int32_t min_orf, Work &w,
		int64_t *n_run)
{
	*n_run = 0;
	if (len < 3) return 1;
	OrfArgs a = {};
	const int64_t nblk = layout(a, len, min_orf);
	const size_t need[6] = { (size_t)nblk * 16 + 16, 16, (size_t)a.cap * 8 + 8, (size_t)a.cap * 4 + 8,
		(size_t)a.cap * 8 + 8, (size_t)a.cap * 4 + 8 };
	for (int i = 0; i < 6; ++i) {
		if (w.n[i] >= need[i]) continue;
		hipError_t e = hipStreamSynchronize(s);
		if (w.p[i]) (void)hipFree(w.p[i]), w.p[i] = nullptr, w.n[i] = 0;
		if (e == hipSuccess) e = hipMalloc(&w.p[i], need[i]);
		if (e != hipSuccess) { nsh_set_error("orf runs: %s", hipGetErrorString(e)); return 0; }
		w.n[i] = need[i];
	}
	a.nt = d_nt, a.codon = d_codon;
	a.lastbad = (int64_t*)w.p[0], a.carry = (int64_t*)w.p[0] + nblk, a.count = (unsigned long long*)w.p[1];
	a.st = (int64_t*)w.p[2], a.n_aa = (int32_t*)w.p[3], a.key = (uint64_t*)w.p[4], a.idx = (uint32_t*)w.p[5];
	w.st = a.st, w.len = a.n_aa;
	hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, s, a.count);
	hipLaunchKernelGGL(orf_summary_kernel, dim3((unsigned)nblk), dim3(NT), 0, s, a);
	hipLaunchKernelGGL(orf_carry_kernel, dim3(3), dim3(NT), 0, s, a);
	hipLaunchKernelGGL(orf_emit_kernel, dim3((unsigned)nblk), dim3(NT), 0, s, a);
	unsigned long long count = 0;
	hipError_t e = hipGetLastError();
	if (e == hipSuccess) e = hipMemcpyAsync(&count, a.count, 8, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (e != hipSuccess) { nsh_set_error("orf runs: %s", hipGetErrorString(e)); return 0; }
	if ((int64_t)count > a.cap) { nsh_set_error("orf runs: %llu runs exceed the bound %lld", count, (long long)a.cap); return 0; }
	nsh_count_launch(4);
	*n_run = (int64_t)count;
	return 1;
}

#define OSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("orf_scan: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" int64_t nsh_orf_scan(const uint8_t *nt, int64_t len, int32_t min_orf, int64_t max_orf, int64_t *orf_st,
		int32_t *orf_len, int64_t max_aa, uint8_t *aa, int64_t *n_aa_total)
{
	if (!nsh_ensure_device()) return -1;
	if (len < 0 || (len > 0 && !nt) || min_orf < 1 || max_orf < 0 || max_aa < 0 || (max_orf > 0 && (!orf_st || !orf_len)) ||
			(max_aa > 0 && !aa) || !n_aa_total) {
		nsh_set_error("orf_scan: bad arguments (min_orf must be >= 1)");
		return -1;
	}
	*n_aa_total = 0;
	if (len < 3) return 0;
	OrfArgs a = {};
	const int64_t nblk = layout(a, len, min_orf);
	std::lock_guard<std::mutex> lk(g_mu);
	static hipStream_t stream = nullptr;
	uint8_t *d_nt = nullptr, *d_codon = nullptr, *d_aa = nullptr;
	void *d_tmp = nullptr;
	int64_t *d_chunk = nullptr, *d_st = nullptr, *d_aaoff = nullptr, *d_st_out = nullptr;
	uint64_t *d_key = nullptr, *d_key2 = nullptr;
	uint32_t *d_idx = nullptr, *d_idx2 = nullptr;
	int32_t *d_len = nullptr, *d_len_out = nullptr;
	unsigned long long *d_count = nullptr, count = 0;
	int64_t n = -1, total = 0;
	size_t tmp_bytes = 0;
	const auto t0 = std::chrono::steady_clock::now();
	{
		if (!stream) OSOK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		OSOK(hipMalloc((void**)&d_nt, (size_t)len + 8));
		OSOK(hipMalloc((void**)&d_codon, 64));
		OSOK(hipMalloc((void**)&d_chunk, (size_t)nblk * 2 * 8));
		OSOK(hipMalloc((void**)&d_count, 8));
		OSOK(hipMalloc((void**)&d_key, (size_t)a.cap * 8));
		OSOK(hipMalloc((void**)&d_key2, (size_t)a.cap * 8));
		OSOK(hipMalloc((void**)&d_idx, (size_t)a.cap * 4));
		OSOK(hipMalloc((void**)&d_idx2, (size_t)a.cap * 4));
		OSOK(hipMalloc((void**)&d_st, (size_t)a.cap * 8));
		OSOK(hipMalloc((void**)&d_len, (size_t)a.cap * 4));
		OSOK(hipMemcpyAsync(d_nt, nt, (size_t)len, hipMemcpyHostToDevice, stream));
		OSOK(hipMemcpyAsync(d_codon, ns_tab_codon, 64, hipMemcpyHostToDevice, stream));
		OSOK(hipMemsetAsync(d_count, 0, 8, stream));
		a.nt = d_nt, a.codon = d_codon, a.lastbad = d_chunk, a.carry = d_chunk + nblk, a.count = d_count;
		a.key = d_key, a.idx = d_idx, a.st = d_st, a.n_aa = d_len;
		hipLaunchKernelGGL(orf_summary_kernel, dim3((unsigned)nblk), dim3(NT), 0, stream, a);
		hipLaunchKernelGGL(orf_carry_kernel, dim3(3), dim3(NT), 0, stream, a);
		hipLaunchKernelGGL(orf_emit_kernel, dim3((unsigned)nblk), dim3(NT), 0, stream, a);
		OSOK(hipGetLastError());
		OSOK(hipMemcpyAsync(&count, d_count, 8, hipMemcpyDeviceToHost, stream));
		OSOK(hipStreamSynchronize(stream));
		if ((int64_t)count > a.cap) { nsh_set_error("orf_scan: %llu runs exceed the bound %lld", count, (long long)a.cap); goto fail; }
		n = (int64_t)count;
		if (n > 0) {
			OSOK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_key, d_key2, d_idx, d_idx2, (int)n, 0, 64, stream));
			OSOK(hipMalloc((void**)&d_aaoff, (size_t)n * 8 + 8));
			OSOK(hipMalloc(&d_tmp, tmp_bytes));
			OSOK(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_key, d_key2, d_idx, d_idx2, (int)n, 0, 64, stream));
			// residue offsets: exclusive sum of the codon counts in sorted order (host bookkeeping over
			// one int per run)
			std::vector<int32_t> len_h(n);
			std::vector<uint32_t> ord(n);
			OSOK(hipMemcpyAsync(ord.data(), d_idx2, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
			OSOK(hipMemcpyAsync(len_h.data(), d_len, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
			OSOK(hipStreamSynchronize(stream));
			std::vector<int64_t> off(n + 1, 0);
			for (int64_t r = 0; r < n; ++r) off[r + 1] = off[r] + len_h[ord[r]];
			total = off[n];
			OSOK(hipMemcpyAsync(d_aaoff, off.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, stream));
			OSOK(hipMalloc((void**)&d_aa, (size_t)total + 8));
			OSOK(hipMalloc((void**)&d_st_out, (size_t)n * 8));
			OSOK(hipMalloc((void**)&d_len_out, (size_t)n * 4));
			hipLaunchKernelGGL(orf_gather_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, stream, a, d_idx2, n,
					d_aaoff, d_st_out, d_len_out, d_aa);
			OSOK(hipGetLastError());
			if (max_orf > 0) {
				OSOK(hipMemcpyAsync(orf_st, d_st_out, (size_t)std::min(n, max_orf) * 8, hipMemcpyDeviceToHost, stream));
				OSOK(hipMemcpyAsync(orf_len, d_len_out, (size_t)std::min(n, max_orf) * 4, hipMemcpyDeviceToHost, stream));
			}
			if (max_aa > 0 && total > 0)
				OSOK(hipMemcpyAsync(aa, d_aa, (size_t)std::min(total, max_aa), hipMemcpyDeviceToHost, stream));
			OSOK(hipStreamSynchronize(stream));
		}
		*n_aa_total = total;
		nsh_count_launch(4);
	}
fail:
	{
		void *bufs[] = { d_nt, d_codon, d_aa, d_tmp, d_chunk, d_st, d_aaoff, d_st_out, d_key, d_key2, d_idx, d_idx2,
			d_len, d_len_out, d_count };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return n;
}
