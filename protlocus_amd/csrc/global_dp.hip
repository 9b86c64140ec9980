// Global spliced DP engine (ns_global_gs16b / ns_global_gs32b with NS_F_CIGAR, the int32
// score-only DP, ns_gpu_submit / ns_gpu_wait; reference semantics nasw-sse.c:351-567, 779-973,
// backtrack 44-93; the reference's batching front end nasw-cuda.cu:419-455, 827-1066).
//
// Requests are queued (inputs copied at submission) and run in batches by one executor thread:
//   * batching: the executor takes everything queued once a caller waits on a queued request, or
//     after a 2-ms window, or at 4096 requests; concurrent callers (kt_for workers) therefore share
//     batches, and a submit-all-then-wait caller (mp_align_gap) gets one batch for all its gaps;
//   * a batch is grouped by options (gap, frameshift, splice and matrix terms; io, the PSSM rows
//     and the strand are per request), then cut so that its trace fits the trace budget;
//   * inputs travel as one block through a pinned staging buffer; the device preparation kernel
//     writes the row records; problems are sorted into buckets of equal shape (waves per
//     work-group, columns per lane; global_dp_impl.h) launched on four streams;
//   * the trace (one uint16 word per cell, the reference's bit layout) stays in HBM: a backtrack
//     kernel walks it, one wave per problem, writes each CIGAR (reversed, tiny U/V turned into G,
//     nasw-sse.c:34-42) at an offset reserved with one atomic, and only the CIGARs cross PCIe;
//   * device and pinned buffers are grow-only pools owned by the executor, so a batch allocates
//     nothing once the pools have reached the workload's size.
// The submitting thread copies the CIGAR into the caller's km (kalloc arenas are per thread).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>
#include "nsh_internal.h"
#include "prep.h"
#include "global_dp_types.h"

using namespace nsh_glob;

namespace {

// ---------------------------------------------------------------------------------------
// device backtrack: one wave per problem. The walk is wave-uniform; its trace words come from a
// window the 64 lanes fetch together, one word each along the last move's direction (an intron is
// a straight run up one column, an exon a 3:1 diagonal), so one memory round trip serves up to
// 64 steps, and each step reads its word from the window lane.
// ---------------------------------------------------------------------------------------
struct BtProb {
	int64_t tr_off, tmp_off;   // trace words; scratch for the ops in push order
	int32_t nl, al;
};

// the ops in push order go to scratch (written by lane 0); the op being extended stays in
// registers (cur, one uint32 len << 4 | op)
struct BtRun {
	uint32_t *t;
	int32_t n;       // ops written to t
	uint32_t cur;    // the open op (0xffffffff: none)
	bool writer;     // lane 0
	__device__ void put(uint32_t c) { if (writer) t[n] = c; ++n; }
	__device__ void push(uint32_t op, int32_t len)
	{
		// ns_push_cigar (nasw.h:219-229): F and G are never merged
		if (cur != 0xffffffffu && op == (cur & 0xf) && op != NS_CIGAR_F && op != NS_CIGAR_G) cur += (uint32_t)len << 4;
		else {
			if (cur != 0xffffffffu) put(cur);
			cur = (uint32_t)len << 4 | op;
		}
	}
	__device__ void close() { if (cur != 0xffffffffu) put(cur), cur = 0xffffffffu; }
};

// grid = problems, 64 threads each
__global__ void __launch_bounds__(64) backtrack_kernel(const BtProb *bp, int32_t n, const uint16_t *trace,
		uint32_t *tmp, uint32_t *packed, unsigned long long *total, int64_t *meta)
{
	const int k = blockIdx.x, L = threadIdx.x;
	if (k >= n) return;
	const BtProb P = bp[k];
	const uint16_t *tr = trace + P.tr_off;
	BtRun b = { tmp + P.tmp_off, 0, 0xffffffffu, L == 0 };
	int32_t i = P.nl - 1, j = P.al - 1, last = 0;
	const int64_t al = P.al;
	// the window: lane q holds the word at (wi - wdi * q, wj - wdj * q); (di, dj): the last move
	int32_t win = 0, wi = -1, wj = -1, wdi = 1, wdj = 0, di = 1, dj = 0;
	while (i >= 2 && j >= 0) {   // nasw-sse.c:44-93
		int32_t t = wdj? wj - j : (wi - i) / wdi;
		if (wj < 0 || t < 0 || t >= 64 || i != wi - wdi * t || j != wj - wdj * t) {
			wi = i, wj = j, wdi = di, wdj = dj, t = 0;
			const int32_t ii = i - di * L, jj = j - dj * L;
			win = ii >= 0 && jj >= 0? tr[(int64_t)ii * al + jj] : 0;
		}
		int32_t x = __builtin_amdgcn_readlane(win, __builtin_amdgcn_readfirstlane(t)), state, ext;
		const int32_t i0 = i, j0 = j;
		if (x >> 9 & 1) x = 1 | x >> 4 << 4;
		state = last == 0? x & 0xf : last;
		ext = state >= 1 && state <= 5? x >> (state + 3) & 1 : 0;
		switch (state) {
		case 0: b.push(NS_CIGAR_M, 1), i -= 3, --j; break;
		case 1: b.push(NS_CIGAR_I, 1), --j; break;
		case 2: b.push(NS_CIGAR_D, 1), i -= 3; break;
		case 3: b.push(NS_CIGAR_N, 1), --i; break;
		case 4: b.push(NS_CIGAR_U, 1), --i; if (!ext) --j; break;
		case 5: b.push(NS_CIGAR_V, 1), --i; if (!ext) --j; break;
		case 6: b.push(NS_CIGAR_F, 1), --i; break;
		case 7: b.push(NS_CIGAR_F, 2), i -= 2; break;
		case 8: b.push(NS_CIGAR_G, 1), --i, --j; break;
		case 9: b.push(NS_CIGAR_G, 2), i -= 2, --j; break;
		default: i = -1; break;   // not produced by the kernels
		}
		last = state >= 1 && state <= 5 && ext? state : 0;
		if (i0 - i > 0 || j0 - j > 0) di = i0 - i, dj = j0 - j;   // every step moves up, left or both
	}
	if (j > 0) b.push(NS_CIGAR_I, j);
	if (i >= 0) {
		const int32_t l = (i + 1) / 3, r = (i + 1) % 3;
		if (l > 0) b.push(NS_CIGAR_D, l);
		if (r != 0) b.push(NS_CIGAR_F, r);
	}
	b.close();
	const int32_t nc = b.n;
	unsigned long long off = 0;
	if (L == 0) off = atomicAdd(total, (unsigned long long)nc);
	off = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(off >> 32)) << 32) |
		(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)off);
	__syncthreads();   // lane 0's scratch writes are visible to the wave
	const uint32_t *t = b.t;
	for (int32_t q = L; q < nc; q += 64) {   // reversed, U/V shorter than 3 become G (nasw-sse.c:34-42)
		uint32_t c = t[nc - 1 - q];
		const uint32_t op = c & 0xf;
		if ((op == NS_CIGAR_U || op == NS_CIGAR_V) && c >> 4 < 3) c = c >> 4 << 4 | NS_CIGAR_G;
		packed[off + q] = c;
	}
	if (L == 0) meta[2 * k] = (int64_t)off, meta[2 * k + 1] = nc;
}

__global__ void fill_u64_zero(unsigned long long *p) { if (threadIdx.x == 0) *p = 0; }

#define HIPOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("%s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

struct DevPool {   // grow-only device buffer
	void *p = nullptr;
	size_t cap = 0;
	hipError_t ensure(size_t n)
	{
		if (n <= cap) return hipSuccess;
		if (p) (void)hipFree(p);
		p = nullptr, cap = 0;
		n = std::max(n + (n >> 3), (size_t)1 << 20);
		const hipError_t e = hipMalloc(&p, n);
		if (e == hipSuccess) cap = n;
		return e;
	}
	template <typename T> T *as(size_t byte_off = 0) const { return (T*)((char*)p + byte_off); }
};
struct HostPool {  // grow-only pinned host buffer
	void *p = nullptr;
	size_t cap = 0;
	hipError_t ensure(size_t n)
	{
		if (n <= cap) return hipSuccess;
		if (p) (void)hipHostFree(p);
		p = nullptr, cap = 0;
		n = std::max(n + (n >> 3), (size_t)1 << 20);
		const hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
		if (e == hipSuccess) cap = n;
		return e;
	}
};

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace

// ---------------------------------------------------------------------------------------
// requests
// ---------------------------------------------------------------------------------------
struct nsh_glob_req {
	std::vector<uint8_t> ns, as, ss;
	std::vector<int8_t> pssm;
	int32_t nl = 0, al = 0, is32 = 0, trace = 0;
	ns_opt_t opt;                // sc/nt4/aa20/codon point into the copies below
	int8_t mat[22 * 22];
	uint8_t nt4[256], aa20[256], codon[64];
	// results
	int done = 0, ok = 0;
	int32_t score = 0;
	std::vector<uint32_t> cigar;
};

namespace {

// one problem as a batch run sees it (inputs borrowed for the run)
struct GlobIn {
	const uint8_t *ns, *as, *ss;
	const int8_t *pssm;
	int64_t g_off;             // >= 0: the bases are in HBM at genome + g_off (ns unused)
	int32_t nl, al, is32, trace, left, io;
};
// what every problem of a run shares (same_terms)
struct GlobTerms {
	int32_t go, ge, fs, sp_null, sp[6];
	int8_t mat[484];
	uint8_t nt4[256], aa20[256], codon[64];
	const uint8_t *genome;     // the oriented genome in HBM that problems with g_off >= 0 read
};

GlobTerms terms_of(const ns_opt_t &o, const uint8_t *genome)
{
	GlobTerms t;
	t.go = o.go, t.ge = o.ge, t.fs = o.fs, t.sp_null = o.sp_null_bonus;
	for (int k = 0; k < 6; ++k) t.sp[k] = o.sp[k];
	std::memcpy(t.mat, o.sc, 484), std::memcpy(t.nt4, o.nt4, 256), std::memcpy(t.aa20, o.aa20, 256);
	std::memcpy(t.codon, o.codon, 64);
	t.genome = genome;
	return t;
}

int kernel_variant(const GlobIn &q) { return q.is32? (q.trace? 1 : 2) : 0; }   // trace16, trace32, score32

// the streams, events and grow-only pools of one thread that runs batches
struct RunCtx {
	hipStream_t st[4] = {};
	hipEvent_t ev_prep = nullptr, ev_b[4] = {};
	DevPool d_in, d_rec, d_aas, d_tr, d_tmp, d_packed, d_meta;
	HostPool h_in, h_meta, h_packed;
	int init()
	{
		if (st[0]) return 1;
		for (int k = 0; k < 4; ++k)
			if (hipStreamCreateWithFlags(&st[k], hipStreamNonBlocking) != hipSuccess ||
					hipEventCreateWithFlags(&ev_b[k], hipEventDisableTiming) != hipSuccess) {
				nsh_set_error("global DP engine: stream set-up failed");
				return 0;
			}
		(void)hipEventCreateWithFlags(&ev_prep, hipEventDisableTiming);
		return 1;
	}
};

// trace bytes per batch part: NSH_TRACE_BUDGET_MB (read per batch), 16 GiB by default
size_t trace_budget()
{
	const char *s = std::getenv("NSH_TRACE_BUDGET_MB");
	return (size_t)(s? std::max(1L, std::atol(s)) : 16384L) << 20;
}

// One batch part on the device: prep, the DP in buckets of equal shape on four streams, the
// backtrack; score[k] and, for problems with a trace, their CIGAR: cig_n[k] ops at cig_off[k] of
// the packed array pk (valid until the context's next run). 1 on success.
int run_list(RunCtx &c, const GlobIn *P, int32_t n, const GlobTerms &T, int32_t *score, int64_t *cig_off,
		int32_t *cig_n, const uint32_t **pk_out)
{
	const auto t0 = std::chrono::steady_clock::now();
	nsh_step_clock clk("global DP");
	// layout of the input block: ns | as | ss | pssm | PrepProb | GlobProb | BtProb | PrepConst | mat
	// Pass 1, one thread: each problem's offsets into the packed arrays, its bucket and its
	// backtrack slot. Buckets are (kernel variant, waves desc, columns per lane desc); a counting
	// sort gives every problem its place in bucket order, problems in input order inside a bucket.
	struct Off { int64_t ns, as, rec, pssm, tr, tmp; };
	std::vector<Off> off(n);
	std::vector<uint16_t> key(n);
	std::vector<int32_t> pos(n), bt_slot(n, -1), bt_req;   // bt_req: problem index of each backtrack slot
	int64_t l_raw = 0, l_rec = 0, l_as = 0, l_pssm = 0, l_tr = 0, l_tmp = 0;
	int any_ss = 0;
	constexpr int kKeys = 3 * (kNWMax + 1) * (kCMaxG + 1);
	int32_t kstart[kKeys + 1] = {};
	for (int32_t k = 0; k < n; ++k) {
		const GlobIn &q = P[k];
		Off &o = off[k];
		o.ns = q.g_off >= 0? q.g_off : l_raw;
		if (q.g_off < 0) l_raw += q.nl + 1;
		o.as = l_as, o.rec = l_rec, o.pssm = q.pssm? l_pssm : -1, o.tr = q.trace? l_tr : 0, o.tmp = l_tmp;
		l_rec += q.nl + 1, l_as += q.al + 1;
		if (q.pssm) l_pssm += (int64_t)q.al * 22;
		if (q.trace) {
			bt_slot[k] = (int32_t)bt_req.size(), bt_req.push_back(k);
			l_tr += (int64_t)q.nl * q.al, l_tmp += (int64_t)q.nl + q.al + 8;
		}
		any_ss |= q.ss != nullptr;
		int C, NW;
		shape(q.al, C, NW);
		if (C > kCMaxG || NW > kNWMax) { nsh_set_error("global DP: protein length %d exceeds %d", q.al, 64 * kCMaxG * kNWMax); return 0; }
		key[k] = (uint16_t)((kernel_variant(q) * (kNWMax + 1) + (kNWMax - NW)) * (kCMaxG + 1) + (kCMaxG - C));
		++kstart[key[k] + 1];
	}
	for (int t = 0; t < kKeys; ++t) kstart[t + 1] += kstart[t];
	{   // inside a bucket, problems of kLongRowsG rows or more first, longest first (they set the
		// bucket's end), the others in input order
		int32_t next[kKeys], nlong[kKeys + 1] = {};
		std::vector<int32_t> longs;
		for (int32_t k = 0; k < n; ++k) if (P[k].nl >= kLongRowsG) ++nlong[key[k]], longs.push_back(k);
		std::stable_sort(longs.begin(), longs.end(), [&](int32_t x, int32_t y) { return P[x].nl > P[y].nl; });
		std::memcpy(next, kstart, sizeof(next));
		for (int32_t k : longs) pos[k] = next[key[k]]++;
		for (int t = 0; t < kKeys; ++t) next[t] = kstart[t] + nlong[t];
		for (int32_t k = 0; k < n; ++k) if (P[k].nl < kLongRowsG) pos[k] = next[key[k]]++;
	}
	// a bucket's kernel variant, waves and columns per lane from its key
	auto key_v = [&](int t) { return t / ((kNWMax + 1) * (kCMaxG + 1)); };
	auto key_nw = [&](int t) { return kNWMax - t / (kCMaxG + 1) % (kNWMax + 1); };
	auto key_c = [&](int t) { return kCMaxG - t % (kCMaxG + 1); };
	if (clk.on) {
		// the batch's shape: per bucket (waves, columns per lane) its problems, rows and cells
		int64_t rows = 0, cells = 0, nl_max = 0, hn[6] = {}, hr[6] = {}, br[kKeys] = {}, bc[kKeys] = {}, bm[kKeys] = {};
		for (int32_t k = 0; k < n; ++k) {
			const int64_t nl = P[k].nl;
			rows += nl, cells += nl * P[k].al, nl_max = std::max(nl_max, nl);
			const int b = nl <= 64? 0 : nl <= 256? 1 : nl <= 1024? 2 : nl <= 4096? 3 : nl <= 16384? 4 : 5;
			++hn[b], hr[b] += nl;
			br[key[k]] += nl, bc[key[k]] += nl * P[k].al, bm[key[k]] = std::max(bm[key[k]], nl);
		}
		std::fprintf(stderr, "[global DP]   %d problems, %lld rows, %.3g cells, longest %lld rows\n", n, (long long)rows,
				(double)cells, (long long)nl_max);
		std::fprintf(stderr, "[global DP]   rows <=64/256/1024/4096/16384/more: problems %lld %lld %lld %lld %lld %lld, rows %lld %lld %lld %lld %lld %lld\n",
				(long long)hn[0], (long long)hn[1], (long long)hn[2], (long long)hn[3], (long long)hn[4], (long long)hn[5],
				(long long)hr[0], (long long)hr[1], (long long)hr[2], (long long)hr[3], (long long)hr[4], (long long)hr[5]);
		for (int t = 0; t < kKeys; ++t) {
			const int32_t m = kstart[t + 1] - kstart[t];
			if (m) std::fprintf(stderr, "[global DP]   bucket v%d NW %d C %d: %d problems, %lld rows (mean %.0f, max %lld), %.3g cells\n",
					key_v(t), key_nw(t), key_c(t), m, (long long)br[t], (double)br[t] / m, (long long)bm[t], (double)bc[t]);
		}
		clk.lap("  offsets and buckets", n);
	}
	const size_t n_bt = bt_req.size();
	const size_t o_ns = 0, o_as = al256(o_ns + l_raw), o_ss = al256(o_as + l_as), o_pssm = al256(o_ss + (any_ss? l_raw : 0));
	const size_t o_pp = al256(o_pssm + l_pssm), o_gp = al256(o_pp + (size_t)n * sizeof(nsh_prep::PrepProb));
	const size_t o_bt = al256(o_gp + (size_t)n * sizeof(GlobProb)), o_pc = al256(o_bt + n_bt * sizeof(BtProb));
	const size_t o_mat = al256(o_pc + sizeof(nsh_prep::PrepConst)), o_pen = al256(o_mat + 484);
	const size_t in_bytes = al256(o_pen + 16);
	int ok = 0;
	uint8_t *h = nullptr, *d = nullptr;
	int32_t *d_score = nullptr;
	int64_t *bmeta = nullptr, *d_btmeta = nullptr;
	unsigned long long total = 0, *d_total = nullptr;
	// device result block: scores (n ints) | CIGAR total | (offset, count) per backtracked problem
	const size_t meta_bytes = al256((size_t)n * 4) + 8 + n_bt * 16;
	*pk_out = nullptr;
	HIPOK(c.h_in.ensure(in_bytes));
	HIPOK(c.d_in.ensure(in_bytes));
	HIPOK(c.d_rec.ensure((size_t)l_rec * 4 + 256));
	HIPOK(c.d_aas.ensure((size_t)l_as + 256));
	if (l_tr) HIPOK(c.d_tr.ensure((size_t)l_tr * 2 + 256));
	if (l_tmp) { HIPOK(c.d_tmp.ensure((size_t)l_tmp * 4 + 256)); HIPOK(c.d_packed.ensure((size_t)l_tmp * 4 + 256)); }
	HIPOK(c.d_meta.ensure(meta_bytes + 256));
	HIPOK(c.h_meta.ensure(meta_bytes + 256));
	clk.lap("  pools", (int64_t)in_bytes);
	h = (uint8_t*)c.h_in.p, d = (uint8_t*)c.d_in.p;
	// pass 2, on the host threads: problem records and strings straight into the pinned block
	{
		nsh_prep::PrepConst pc;
		std::memcpy(pc.nt4, T.nt4, 256), std::memcpy(pc.aa20, T.aa20, 256), std::memcpy(pc.codon, T.codon, 64);
		for (int t = 0; t < 6; ++t) pc.sp[t] = T.sp[t];
		pc.sp_null = T.sp_null;
		pc.genome = T.genome;
		nsh_prep::PrepProb *hpp = (nsh_prep::PrepProb*)(h + o_pp);
		GlobProb *hgp = (GlobProb*)(h + o_gp);
		BtProb *hbt = (BtProb*)(h + o_bt);
		const int32_t nth = (int32_t)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
		nsh_parallel_for((n + 4095) / 4096, nth, [&](int64_t b) {
			for (int32_t k = (int32_t)(b * 4096); k < std::min<int64_t>(n, (b + 1) * 4096); ++k) {
				const GlobIn &q = P[k];
				const Off &o = off[k];
				nsh_prep::PrepProb p = {};
				p.from_genome = q.g_off >= 0, p.ns_off = o.ns;
				p.as_off = o.as, p.rec_off = o.rec, p.aa_off = o.as;
				// the strand only matters to int32 score-only requests with an extension flag (nasw-sse.c:246-247)
				p.nl = q.nl, p.al = q.al, p.left = q.left, p.io = q.io, p.has_ss = q.ss != nullptr;
				hpp[k] = p;
				GlobProb g = {};
				g.rec_off = o.rec, g.aa_off = o.as, g.pssm_off = o.pssm, g.tr_off = o.tr;
				g.nl = q.nl, g.al = q.al, g.io = q.io, g.out = k, g.left = q.left;
				g.C = key_c(key[k]), g.NW = key_nw(key[k]);
				const int vs = q.is32? 4 : 8;
				g.slen = (q.al + vs - 1) / vs, g.npad = q.al, g.mode = 0;
				hgp[pos[k]] = g;
				if (bt_slot[k] >= 0) hbt[bt_slot[k]] = { o.tr, o.tmp, q.nl, q.al };
				if (q.nl && !p.from_genome) std::memcpy(h + o_ns + o.ns, q.ns, q.nl);
				std::memcpy(h + o_as + o.as, q.as, q.al);
				if (any_ss && !p.from_genome) {
					if (q.ss) std::memcpy(h + o_ss + o.ns, q.ss, q.nl);
					else std::memset(h + o_ss + o.ns, 0xff, q.nl);
				}
				if (q.pssm) std::memcpy(h + o_pssm + o.pssm, q.pssm, (size_t)q.al * 22);
			}
		});
		std::memcpy(h + o_pc, &pc, sizeof(pc));
		std::memcpy(h + o_mat, T.mat, 484);
	}
	clk.lap("  pack", n);
	{
		const auto t1 = std::chrono::steady_clock::now();
		HIPOK(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, c.st[0]));
		nsh_add_times(std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count(), 0, 0);
	}
	d_score = c.d_meta.as<int32_t>();
	nsh_prep::launch((const nsh_prep::PrepProb*)(d + o_pp), n, d + o_ns, any_ss? d + o_ss : nullptr, d + o_as,
			c.d_rec.as<uint32_t>(), c.d_aas.as<uint8_t>(), (const nsh_prep::PrepConst*)(d + o_pc), c.st[0]);
	HIPOK(hipGetLastError());
	HIPOK(hipEventRecord(c.ev_prep, c.st[0]));
	if (clk.on) {
		HIPOK(hipStreamSynchronize(c.st[0]));
		clk.lap("  upload, prep", n);
	}
	{
		GlobArgs ga = {};
		ga.rec = c.d_rec.as<uint32_t>(), ga.aas = c.d_aas.as<uint8_t>(), ga.pssm = (const int8_t*)(d + o_pssm);
		ga.mat = (const int8_t*)(d + o_mat), ga.asize = 22;
		ga.go = T.go, ga.ge = T.ge, ga.fs = T.fs;
		ga.trace = l_tr? c.d_tr.as<uint16_t>() : nullptr, ga.score = d_score;
		ga.pen = (const int32_t*)(d + o_pen), ga.pen_n = 1;
		const GlobProb *d_gp = (const GlobProb*)(d + o_gp);
		int nb = 0;
		for (int t = 0; t < kKeys; ++t) {
			const int32_t k = kstart[t], m = kstart[t + 1] - k;
			if (!m) continue;
			hipStream_t s = c.st[nb & 3];
			HIPOK(hipStreamWaitEvent(s, c.ev_prep, 0));
			ga.prob = d_gp + k;
			const int v = key_v(t);
			if (v == 0) launch_trace16(ga, m, key_nw(t), key_c(t), s);
			else if (v == 1) launch_trace32(ga, m, key_nw(t), key_c(t), s);
			else launch_score32(ga, m, key_nw(t), key_c(t), s);
			HIPOK(hipGetLastError());
			++nb;
		}
		for (int b = 1; b < std::min(nb, 4); ++b) {
			HIPOK(hipEventRecord(c.ev_b[b], c.st[b]));
			HIPOK(hipStreamWaitEvent(c.st[0], c.ev_b[b], 0));
		}
		nsh_count_launch(1 + nb);
		if (clk.on) {
			HIPOK(hipStreamSynchronize(c.st[0]));
			clk.lap("  DP kernels", nb);
		}
	}
	d_btmeta = (int64_t*)((char*)c.d_meta.p + al256((size_t)n * 4) + 8);
	d_total = (unsigned long long*)((char*)c.d_meta.p + al256((size_t)n * 4));
	if (n_bt) {
		hipLaunchKernelGGL(fill_u64_zero, dim3(1), dim3(64), 0, c.st[0], d_total);
		hipLaunchKernelGGL(backtrack_kernel, dim3((int)n_bt), dim3(64), 0, c.st[0], (const BtProb*)(d + o_bt),
				(int32_t)n_bt, c.d_tr.as<uint16_t>(), c.d_tmp.as<uint32_t>(), c.d_packed.as<uint32_t>(), d_total, d_btmeta);
		HIPOK(hipGetLastError());
		nsh_count_launch(2);
	}
	// scores, backtrack counts and offsets, then the CIGARs
	HIPOK(hipMemcpyAsync(c.h_meta.p, c.d_meta.p, meta_bytes, hipMemcpyDeviceToHost, c.st[0]));
	HIPOK(hipStreamSynchronize(c.st[0]));
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	clk.lap("  backtrack, results", (int64_t)n_bt);
	bmeta = (int64_t*)((char*)c.h_meta.p + al256((size_t)n * 4) + 8);
	total = !n_bt? 0 : *(unsigned long long*)((char*)c.h_meta.p + al256((size_t)n * 4));
	if (total) {
		const auto t3 = std::chrono::steady_clock::now();
		HIPOK(c.h_packed.ensure((size_t)total * 4));
		HIPOK(hipMemcpyAsync(c.h_packed.p, c.d_packed.p, (size_t)total * 4, hipMemcpyDeviceToHost, c.st[0]));
		HIPOK(hipStreamSynchronize(c.st[0]));
		nsh_add_times(0, 0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t3).count());
		clk.lap("  CIGAR download", (int64_t)total);
	}
	{
		const int32_t *hs = (const int32_t*)c.h_meta.p;
		for (int32_t k = 0; k < n; ++k) {
			score[k] = P[k].nl > 2? hs[k] : (P[k].is32? -0x40000000 : -32768);
			cig_off[k] = 0, cig_n[k] = 0;
		}
		for (size_t b = 0; b < n_bt; ++b) cig_off[bt_req[b]] = bmeta[2 * b], cig_n[bt_req[b]] = (int32_t)bmeta[2 * b + 1];
		*pk_out = (const uint32_t*)c.h_packed.p;
	}
	ok = 1;
fail:
	return ok;
}

bool same_terms(const nsh_glob_req *a, const nsh_glob_req *b)
{
	const ns_opt_t &x = a->opt, &y = b->opt;
	return x.go == y.go && x.ge == y.ge && x.fs == y.fs && x.sp_null_bonus == y.sp_null_bonus &&
		std::memcmp(x.sp, y.sp, sizeof(x.sp)) == 0 && std::memcmp(a->mat, b->mat, sizeof(a->mat)) == 0 &&
		std::memcmp(a->nt4, b->nt4, 256) == 0 && std::memcmp(a->aa20, b->aa20, 256) == 0 &&
		std::memcmp(a->codon, b->codon, 64) == 0;
}

class Engine {
public:
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<nsh_glob_req*> queue;
	int n_waiting = 0;
	bool started = false;
	RunCtx ctx;   // executor-owned

	void start()
	{
		if (started) return;
		started = true;
		std::thread([this] { loop(); }).detach();
	}

	void submit(nsh_glob_req *q)
	{
		std::lock_guard<std::mutex> lk(mu);
		start();
		queue.push_back(q);
		cv_work.notify_one();
	}

	void submit_many(const std::vector<nsh_glob_req*> &qs)   // queued together: one batch
	{
		std::lock_guard<std::mutex> lk(mu);
		start();
		for (nsh_glob_req *q : qs) queue.push_back(q);
		++n_waiting;              // the caller waits next: no batching window
		cv_work.notify_one();
	}

	void wait(nsh_glob_req *q)
	{
		std::unique_lock<std::mutex> lk(mu);
		if (q->done) return;
		++n_waiting;
		cv_work.notify_one();   // a waiter ends the batching window
		cv_done.wait(lk, [&] { return q->done != 0; });
		--n_waiting;
	}

	void loop()
	{
		for (;;) {
			std::vector<nsh_glob_req*> batch;
			{
				std::unique_lock<std::mutex> lk(mu);
				cv_work.wait(lk, [&] { return !queue.empty(); });
				// batching window: until a caller waits, 4096 requests are queued, or 2 ms pass
				const auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(2);
				while (n_waiting == 0 && queue.size() < 4096 &&
						cv_work.wait_until(lk, until) != std::cv_status::timeout) {}
				batch.assign(queue.begin(), queue.end());
				queue.clear();
			}
			run(batch);
			{
				std::lock_guard<std::mutex> lk(mu);
				for (nsh_glob_req *q : batch) q->done = 1;
			}
			cv_done.notify_all();
		}
	}

	void run(std::vector<nsh_glob_req*> &batch)
	{
		if (!nsh_ensure_device() || !ctx.init()) { for (auto *q : batch) q->ok = 0; return; }
		// option groups, then trace-budget cuts
		const size_t budget = trace_budget();
		std::vector<char> taken(batch.size(), 0);
		for (size_t a = 0; a < batch.size(); ++a) {
			if (taken[a]) continue;
			std::vector<nsh_glob_req*> grp;
			for (size_t b = a; b < batch.size(); ++b)
				if (!taken[b] && same_terms(batch[a], batch[b])) taken[b] = 1, grp.push_back(batch[b]);
			size_t lo = 0;
			while (lo < grp.size()) {
				size_t hi = lo, tr = 0;
				while (hi < grp.size()) {
					const size_t t = grp[hi]->trace? (size_t)grp[hi]->nl * grp[hi]->al * 2 : 0;
					if (hi > lo && tr + t > budget) break;
					tr += t, ++hi;
				}
				std::vector<nsh_glob_req*> part(grp.begin() + lo, grp.begin() + hi);
				if (!run_part(part)) for (auto *q : part) q->ok = 0;
				lo = hi;
			}
		}
	}

	int run_part(std::vector<nsh_glob_req*> &reqs)
	{
		const int32_t n = (int32_t)reqs.size();
		std::vector<GlobIn> in(n);
		for (int32_t k = 0; k < n; ++k) {
			const nsh_glob_req *q = reqs[k];
			in[k] = { q->ns.data(), q->as.data(), q->ss.empty()? nullptr : q->ss.data(),
				q->pssm.empty()? nullptr : q->pssm.data(), -1, q->nl, q->al, q->is32, q->trace,
				(q->opt.flag & NS_F_EXT_LEFT) != 0, q->opt.io };
		}
		std::vector<int32_t> sc(n), cn(n);
		std::vector<int64_t> co(n);
		const uint32_t *pk = nullptr;
		if (!run_list(ctx, in.data(), n, terms_of(reqs[0]->opt, nullptr), sc.data(), co.data(), cn.data(), &pk)) return 0;
		for (int32_t k = 0; k < n; ++k) {
			nsh_glob_req *q = reqs[k];
			q->score = sc[k], q->ok = 1;
			if (cn[k]) q->cigar.assign(pk + co[k], pk + co[k] + cn[k]);
		}
		return 1;
	}
};

Engine &engine()
{
	static Engine *e = new Engine();   // never destroyed: its thread lives until the process ends
	return *e;
}

// the direct batch path's own context (nsh_global_batch runs on the caller's thread)
std::mutex g_direct_mu;
RunCtx &direct_ctx()
{
	static RunCtx *c = new RunCtx();
	return *c;
}

} // namespace

// ---------------------------------------------------------------------------------------
// internal API (nsh_internal.h)
// ---------------------------------------------------------------------------------------
// a request with its inputs copied (nullptr and the reason in ns_gpu_error() when it is refused)
static nsh_glob_req *glob_make(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t is32)
{
	if (!nsh_ensure_device()) return nullptr;
	const bool cigar = (opt->flag & NS_F_CIGAR) && !(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	if (!is32 && !cigar) { nsh_set_error("score-only int16 DP goes through the extension engine"); return nullptr; }
	if (al < 1 || nl < 0) { nsh_set_error("global DP: nl=%d al=%d", nl, al); return nullptr; }
	if (al > 64 * kCMaxG * kNWMax) { nsh_set_error("global DP: protein length %d exceeds %d", al, 64 * kCMaxG * kNWMax); return nullptr; }
	if (opt->asize != 22) { nsh_set_error("asize %d not supported", opt->asize); return nullptr; }
	if (opt->ge < 0 || opt->go < 0 || opt->ge > 32767 || opt->go > 32767 || opt->fs < 0 || opt->fs > 32767) {
		nsh_set_error("gap penalties go=%d ge=%d fs=%d out of the supported range", opt->go, opt->ge, opt->fs); return nullptr;
	}
	for (int t = 0; t < 64; ++t)
		if (opt->codon[t] >= 22) { nsh_set_error("codon table entry %d out of range", t); return nullptr; }
	for (int t = 0; t < 256; ++t)
		if (opt->aa20[t] >= 22 || opt->nt4[t] > 4) { nsh_set_error("alphabet table entry %d out of range", t); return nullptr; }
	nsh_glob_req *q = new nsh_glob_req();
	q->nl = nl, q->al = al, q->is32 = is32 != 0, q->trace = cigar;
	q->ns.assign((const uint8_t*)ns, (const uint8_t*)ns + nl);
	q->as.assign((const uint8_t*)as, (const uint8_t*)as + al);
	if (ss) q->ss.assign(ss, ss + nl);
	if (opt->pssm) q->pssm.assign(opt->pssm, opt->pssm + (size_t)al * 22);
	q->opt = *opt;
	std::memcpy(q->mat, opt->sc, 484), std::memcpy(q->nt4, opt->nt4, 256), std::memcpy(q->aa20, opt->aa20, 256);
	std::memcpy(q->codon, opt->codon, 64);
	q->opt.sc = q->mat, q->opt.nt4 = q->nt4, q->opt.aa20 = q->aa20, q->opt.codon = q->codon, q->opt.pssm = nullptr;
	return q;
}

nsh_glob_req *nsh_glob_submit(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t is32)
{
	nsh_glob_req *q = glob_make(ns, nl, as, al, opt, ss, is32);
	if (q) engine().submit(q);
	return q;
}

int nsh_glob_wait(nsh_glob_req *q, void *km, ns_rst_t *r)
{
	if (!q) return 0;
	engine().wait(q);
	const int ok = q->ok;
	if (ok) {   // the caller's r may already own a CIGAR buffer (ns_rst_init): append into it
		for (uint32_t c : q->cigar) r->cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, r->cigar, c & 0xf, (int32_t)(c >> 4));
		r->nt_len = q->nl, r->aa_len = q->al, r->score = q->score;
	}
	delete q;
	return ok;
}

int nsh_global_run(void *km, const char *ns, int32_t nl, const char *as, int32_t al,
		const ns_opt_t *opt, const uint8_t *ss, int32_t is32, ns_rst_t *r)
{
	nsh_glob_req *q = nsh_glob_submit(ns, nl, as, al, opt, ss, is32);
	if (!q) return 0;
	r->n_cigar = 0;
	return nsh_glob_wait(q, km, r);
}

// n problems of one option set in one batch, run on the caller's thread (nasw-hip.h); with d_g,
// problem k with g_off[k] >= 0 reads its bases from d_g + g_off[k] in HBM instead of ns
int nsh_global_batch_g(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *ss, const ns_opt_t *opt, int32_t mode,
		const uint8_t *d_g, const int64_t *g_off, int32_t *score, int32_t *n_cigar, uint32_t **cigar)
{
	if (n < 0 || (n > 0 && (!ns_off || !nl || !as || !as_off || !al || (!ns && !g_off) || (g_off && !d_g))) || !opt ||
			!score || !n_cigar || !cigar) {
		nsh_set_error("nsh_global_batch: bad arguments");
		return 0;
	}
	*cigar = nullptr;
	if (n == 0) return 1;
	if (!nsh_ensure_device()) return 0;
	// the checks glob_make makes, once for the shared options and per problem for the lengths
	const bool trace = (opt->flag & NS_F_CIGAR) && !(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	if (opt->asize != 22) { nsh_set_error("asize %d not supported", opt->asize); return 0; }
	if (opt->ge < 0 || opt->go < 0 || opt->ge > 32767 || opt->go > 32767 || opt->fs < 0 || opt->fs > 32767) {
		nsh_set_error("gap penalties go=%d ge=%d fs=%d out of the supported range", opt->go, opt->ge, opt->fs); return 0;
	}
	for (int t = 0; t < 64; ++t)
		if (opt->codon[t] >= 22) { nsh_set_error("codon table entry %d out of range", t); return 0; }
	for (int t = 0; t < 256; ++t)
		if (opt->aa20[t] >= 22 || opt->nt4[t] > 4) { nsh_set_error("alphabet table entry %d out of range", t); return 0; }
	int32_t max_sub = 0;
	for (int32_t t = 0; t < opt->asize * opt->asize; ++t) if (max_sub < opt->sc[t]) max_sub = opt->sc[t];
	std::vector<GlobIn> in(n);
	for (int32_t k = 0; k < n; ++k) {
		if (al[k] < 1 || nl[k] < 0 || al[k] > 64 * kCMaxG * kNWMax) {
			nsh_set_error("nsh_global_batch: problem %d: nl=%d al=%d (protein limit %d)", k, nl[k], al[k], 64 * kCMaxG * kNWMax);
			return 0;
		}
		// mode 2: ns_global_gs_auto's choice (nasw-sse.c:980-999), per problem
		const int32_t is32 = mode == 1 || (mode == 2 && (int64_t)al[k] * max_sub + opt->end_bonus > 30000);
		if (!is32 && !trace) { nsh_set_error("score-only int16 DP goes through the extension engine"); return 0; }
		const bool from_g = g_off && g_off[k] >= 0;
		if (from_g && ss) { nsh_set_error("nsh_global_batch: splice scores need bases on the host"); return 0; }
		in[k] = { from_g? nullptr : (const uint8_t*)ns + ns_off[k], (const uint8_t*)as + as_off[k],
			ss? ss + ns_off[k] : nullptr, opt->pssm, from_g? g_off[k] : -1, nl[k], al[k], is32, trace,
			(opt->flag & NS_F_EXT_LEFT) != 0, opt->io };
	}
	const GlobTerms T = terms_of(*opt, d_g);
	std::vector<int64_t> co(n);
	std::vector<int32_t> cn(n);
	std::vector<uint32_t> all;
	std::lock_guard<std::mutex> lk(g_direct_mu);
	RunCtx &ctx = direct_ctx();
	if (!ctx.init()) return 0;
	// trace-budget cuts over the input order
	const size_t budget = trace_budget();
	for (int32_t lo = 0; lo < n;) {
		int32_t hi = lo;
		size_t tr = 0;
		while (hi < n) {
			const size_t t = trace? (size_t)nl[hi] * al[hi] * 2 : 0;
			if (hi > lo && tr + t > budget) break;
			tr += t, ++hi;
		}
		const uint32_t *pk = nullptr;
		if (!run_list(ctx, in.data() + lo, hi - lo, T, score + lo, co.data() + lo, cn.data() + lo, &pk)) return 0;
		for (int32_t k = lo; k < hi; ++k) {
			const int64_t o = (int64_t)all.size();
			if (cn[k]) all.insert(all.end(), pk + co[k], pk + co[k] + cn[k]);
			co[k] = o;
		}
		lo = hi;
	}
	uint32_t *out = (uint32_t*)std::malloc(std::max<size_t>(all.size(), 1) * sizeof(uint32_t));
	if (!out) { nsh_set_error("nsh_global_batch: out of host memory"); return 0; }
	if (!all.empty()) std::memcpy(out, all.data(), all.size() * sizeof(uint32_t));
	for (int32_t k = 0; k < n; ++k) n_cigar[k] = cn[k];
	*cigar = out;
	return 1;
}

extern "C" int nsh_global_batch(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *ss, const ns_opt_t *opt, int32_t mode,
		int32_t *score, int32_t *n_cigar, uint32_t **cigar)
{
	if (n > 0 && !ns) { nsh_set_error("nsh_global_batch: bad arguments"); return 0; }
	return nsh_global_batch_g(n, ns, ns_off, nl, as, as_off, al, ss, opt, mode, nullptr, nullptr, score, n_cigar, cigar);
}
