// Gapless X-drop candidate scoring (ns_gpu_gapless_*: nasw.h:157-184; semantics of
// nasw-cuda.cu:1151-1295, host side 1296-1412).
//
// One wave64 scores one candidate. For each of its (at most 32) anchors:
//   * seed: the lanes split the block_width target end positions, each lane sums the k-mer
//     score at its positions (codons translated on the fly from the 4-bit packed genome, the
//     reverse strand complemented on the fly), and a wave reduction keeps the best score with
//     the smallest target end;
//   * extension: left and right X-drop walks run 64 steps at a time: every lane scores one step,
//     a wave prefix sum gives the running score, a prefix max the running best, and the first
//     step that leaves the diagonal or drops more than xdrop below the best ends the walk;
//   * the wave keeps the segments in score order (ties: later after earlier) in LDS and, after the
//     last anchor, sums them with the overlap discount sc * (len - ov) / len.
// The genome stays resident in HBM across calls (re-uploaded only when the caller's buffer or
// its length changes, or after ns_gpu_gapless_reset_target()).
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include "nsh_internal.h"

namespace {

#define DEV static __device__ __forceinline__

constexpr int kWaves = 4;          // candidates per work-group
constexpr int kMaxAnchors = 32;
constexpr int kMaxKmer = 16;
constexpr int kMaxAsize = 32;
constexpr int kChunk = 256;        // seed positions staged per pass
constexpr int kWin = kChunk + 3 * (kMaxKmer - 1);

struct GlArgs {
	const uint8_t *pk;                            // 4-bit packed genome, 2 bases per byte (low nibble first)
	const uint8_t *q;                             // aa codes of all queries
	const ns_gpu_gapless_candidate_t *cand;
	const ns_gpu_gapless_anchor_t *anch;
	int32_t n, bw, kmer, xdrop, asize;
	const int8_t *mat;
	const uint8_t *codon;
	int32_t *score;
};

// base at oriented position pos of a contig (4 outside it); reverse: complement of the mirrored base
DEV int gl_base(const uint8_t *pk, int64_t off, int64_t len, int rev, int64_t pos)
{
	if (pos < 0 || pos >= len) return 4;
	const int64_t a = off + (rev? len - 1 - pos : pos);
	const int b = (pk[a >> 1] >> ((a & 1) * 4)) & 0xf;
	return rev && b < 4? 3 - b : b;
}
// amino acid of the codon ending at oriented position end; 20 (stop) for any base outside ACGT
DEV int gl_aa(const uint8_t *pk, const uint8_t *codon, int64_t off, int64_t len, int rev, int64_t end)
{
	const int b0 = gl_base(pk, off, len, rev, end - 2), b1 = gl_base(pk, off, len, rev, end - 1),
		b2 = gl_base(pk, off, len, rev, end);
	return b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 20;
}

template <int CTRL, int ROWS = 0xf>
DEV int dpp(int old, int src) { return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWS, 0xf, false); }

// inclusive wave scans: shifts 1-8 inside each row of 16 lanes, then lane 15 and lane 31 onwards
DEV int wave_sum_incl(int v)
{
	v += dpp<0x111>(0, v);
	v += dpp<0x112>(0, v);
	v += dpp<0x114>(0, v);
	v += dpp<0x118>(0, v);
	v += dpp<0x142, 0xa>(0, v);
	v += dpp<0x143, 0xc>(0, v);
	return v;
}
DEV int wave_max_incl(int v)
{
	v = max(v, dpp<0x111>(INT_MIN, v));
	v = max(v, dpp<0x112>(INT_MIN, v));
	v = max(v, dpp<0x114>(INT_MIN, v));
	v = max(v, dpp<0x118>(INT_MIN, v));
	v = max(v, dpp<0x142, 0xa>(INT_MIN, v));
	v = max(v, dpp<0x143, 0xc>(INT_MIN, v));
	return v;
}

DEV void wave_sync_lds()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one X-drop walk along the diagonal, 64 steps per iteration; step(st, &v) is false once step st
// (1-based) leaves the sequences. Sequential meaning: stop before an invalid step; after a valid
// step add its score, raise the best on a strict gain, and stop when the best exceeds the running
// score by more than xdrop. Returns the best gain (>= 0); *len = the step where it was first reached.
template <class F>
DEV int xdrop_walk(const F &step, int xdrop, int L, int *len)
{
	int lr = 0, lb = 0, ll = 0;
	for (int base = 1;; base += 64) {
		int v = 0;
		const bool valid = step(base + L, &v);
		const int S = lr + wave_sum_incl(valid? v : 0);
		const int M = wave_max_incl(S);                     // best running score over steps <= this one
		const uint64_t sb = __ballot(!valid || max(lb, M) - S > xdrop);
		const int f = sb? __builtin_ctzll(sb) : 64;
		const int np = f == 64? 64 : (((__ballot(valid) >> f) & 1)? f + 1 : f);   // steps taken
		const int cm = np > 0? __builtin_amdgcn_readlane(M, np - 1) : INT_MIN;
		if (cm > lb) {
			ll = base + __builtin_ctzll(__ballot(L < np && S == cm));
			lb = cm;
		}
		if (sb) break;   // the walk ended inside this chunk (possibly after its last step)
		lr = __builtin_amdgcn_readlane(S, 63);
	}
	*len = ll;
	return lb;
}

__global__ void __launch_bounds__(64 * kWaves) gapless_kernel(GlArgs a)
{
	__shared__ int8_t smat[kMaxAsize * kMaxAsize];
	__shared__ uint8_t scod[64];
	__shared__ uint8_t swin[kWaves][kWin];           // amino acids of the staged seed window
	__shared__ int8_t scol[kWaves][kMaxKmer][kMaxAsize];   // per seed position: mat[aa][query residue]
	__shared__ int seg[kWaves][3][kMaxAnchors];
	const int asz = a.asize;
	for (int t = threadIdx.x; t < asz * asz; t += blockDim.x) smat[t] = a.mat[t];
	for (int t = threadIdx.x; t < 64; t += blockDim.x) scod[t] = a.codon[t];
	__syncthreads();
	const int L = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
	const int ci = blockIdx.x * kWaves + w;
	if (ci >= a.n) return;   // whole waves leave; no barrier follows
	const ns_gpu_gapless_candidate_t c = a.cand[ci];
	const uint8_t *qq = a.q + c.query_offset;
	const int kmer = a.kmer, xdrop = a.xdrop, ext = 3 * (kmer - 1);
	const int64_t off = c.target_offset, tlen = c.target_length;
	const int rev = c.reverse;
	const int na = min(c.anchor_count, kMaxAnchors);
	int *sq = seg[w][0], *se = seg[w][1], *sv = seg[w][2];
	uint8_t *win = swin[w];
	int8_t (*col)[kMaxAsize] = scol[w];
	int n_seg = 0;
	for (int an = 0; an < na; ++an) {
		const ns_gpu_gapless_anchor_t A = a.anch[c.anchor_offset + an];
		const int bq = A.query_end;
		if (bq < kmer - 1 || bq >= c.query_length) continue;
		// this anchor's k-mer of the query, as matrix columns
		wave_sync_lds();
		for (int t = L; t < kmer * asz; t += 64) {
			const int s = t / asz, r = t - s * asz;
			col[s][r] = smat[r * asz + qq[bq - kmer + 1 + s]];
		}
		// best k-mer seed over the block; ties go to the smallest target end
		int best = INT_MIN;
		int64_t bt = INT64_MAX;
		for (int d0 = 0; d0 < a.bw; d0 += kChunk) {
			const int nd = min(kChunk, a.bw - d0);
			const int64_t p0 = A.target_block + d0 - ext;   // codon end of window slot 0
			wave_sync_lds();
			for (int t = L; t < nd + ext; t += 64) win[t] = (uint8_t)gl_aa(a.pk, scod, off, tlen, rev, p0 + t);
			wave_sync_lds();
			for (int d = L; d < nd; d += 64) {
				const int64_t te = A.target_block + d0 + d;
				if (te - ext - 2 < 0 || te >= tlen) continue;
				int sc = 0;
				for (int s = 0; s < kmer; ++s) sc += col[s][win[d + 3 * s]];
				if (sc > best) best = sc, bt = te;
			}
		}
#pragma unroll
		for (int d = 32; d >= 1; d >>= 1) {
			const int ob = __shfl_xor(best, d);
			const int64_t ot = __shfl_xor(bt, d);
			if (ob > best || (ob == best && ot < bt)) best = ob, bt = ot;
		}
		if (best == INT_MIN) continue;   // no usable seed: no segment
		int ll, rl;
		const int lb = xdrop_walk([&](int st, int *v) {
			const int qp = bq - kmer - (st - 1);
			const int64_t te = bt - (int64_t)(kmer - 1 + st) * 3;
			if (qp < 0 || te - 2 < 0) return false;
			*v = smat[gl_aa(a.pk, scod, off, tlen, rev, te) * asz + qq[qp]];
			return true;
		}, xdrop, L, &ll);
		const int rb = xdrop_walk([&](int st, int *v) {
			const int qp = bq + st;
			const int64_t te = bt + (int64_t)st * 3;
			if (qp >= c.query_length || te >= tlen) return false;
			*v = smat[gl_aa(a.pk, scod, off, tlen, rev, te) * asz + qq[qp]];
			return true;
		}, xdrop, L, &rl);
		{   // insert into the score-descending list, after equal scores. Every lane runs this
			// wave-uniform code on the same LDS words with the same values (no lane branch).
			const int sc = best + lb + rb, qs = bq - (kmer - 1) - ll, qe = bq + rl;
			int j = n_seg;
			for (; j > 0 && sv[j - 1] < sc; --j) sq[j] = sq[j - 1], se[j] = se[j - 1], sv[j] = sv[j - 1];
			sq[j] = qs, se[j] = qe, sv[j] = sc;
		}
		++n_seg;
	}
	{
		int total = 0;
		for (int i = 0; i < n_seg; ++i) {
			const int len = se[i] - sq[i] + 1;
			if (i > 0 && sv[i] < xdrop) break;
			int ov = 0;
			for (int j = 0; j < i; ++j) {
				const int s = max(sq[i], sq[j]), e = min(se[i], se[j]);
				if (sv[j] != INT_MIN && e >= s) ov += e - s + 1;
			}
			if (ov >= len) { sv[i] = INT_MIN; continue; }
			total += ov > 0? (int)((int64_t)sv[i] * (len - ov) / len) : sv[i];
		}
		a.score[ci] = n_seg? total : INT_MIN;
	}
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
struct DevBuf { void *p = nullptr; size_t cap = 0; };
enum { B_TARGET, B_QUERY, B_CAND, B_ANCH, B_MAT, B_CODON, B_SCORE, B_COUNT };
std::mutex g_mu;   // calls come from worker threads; buffers and the stream are shared
DevBuf g_buf[B_COUNT];
const uint8_t *g_target_src = nullptr;
int64_t g_target_bases = 0;
hipStream_t g_stream = nullptr;
std::atomic<uint64_t> g_batches{0}, g_cands{0}, g_anchors{0}, g_seedpos{0};

bool ensure(DevBuf &b, size_t n)
{
	if (b.cap >= n && b.p) return true;
	if (b.p) (void)hipFree(b.p), b.p = nullptr, b.cap = 0;
	const size_t cap = n + n / 4 + 256;
	if (hipMalloc(&b.p, cap) != hipSuccess) { nsh_set_error("gapless: hipMalloc(%zu) failed", cap); return false; }
	b.cap = cap;
	return true;
}

} // namespace

extern "C" void ns_gpu_gapless_profile(ns_gpu_gapless_profile_t *profile)
{
	profile->batches = g_batches.load(), profile->candidates = g_cands.load();
	profile->anchors = g_anchors.load(), profile->seed_positions = g_seedpos.load();
}

extern "C" int ns_gpu_gapless_supported(void) { return nsh_ensure_device()? 1 : 0; }

extern "C" void ns_gpu_gapless_reset_target(void)
{
	std::lock_guard<std::mutex> lk(g_mu);
	g_target_src = nullptr, g_target_bases = 0;
}

#define GLOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("gapless: %s: %s", #x, hipGetErrorString(e_)); return 0; } } while (0)

namespace {

// the arguments of one scoring batch; every device access stays inside the buffers: residue
// codes index the matrix, contigs lie in the genome, queries and anchors in their arrays
struct ScoreIn {
	const uint8_t *packed_nt; int64_t packed_nt_bases;
	const uint8_t *queries; int64_t query_aa;
	const ns_gpu_gapless_candidate_t *candidates; int32_t n_candidates;
	const ns_gpu_gapless_anchor_t *anchors; int64_t n_anchors;
	int32_t block_width, kmer, xdrop;
	const int8_t *score_matrix; int32_t asize; const uint8_t *codon;
	uint64_t n_anch_used = 0;
};

int check_inputs(ScoreIn &in)
{
	if (!in.packed_nt || in.packed_nt_bases < 1 || !in.queries || in.query_aa < 1 || !in.candidates ||
			in.n_candidates < 1 || !in.anchors || in.n_anchors < 1 || in.block_width < 1 || in.kmer < 1 ||
			in.xdrop < 0 || !in.score_matrix || in.asize < 1 || !in.codon)
		return 0;
	if (!nsh_ensure_device()) return 0;
	if (in.asize < 21 || in.asize > kMaxAsize) {
		nsh_set_error("gapless: asize %d outside 21..%d (stop codons score row 20)", in.asize, kMaxAsize);
		return 0;
	}
	if (in.kmer > kMaxKmer) { nsh_set_error("gapless: kmer %d > %d", in.kmer, kMaxKmer); return 0; }
	for (int t = 0; t < 64; ++t) if (in.codon[t] >= in.asize) { nsh_set_error("gapless: codon table out of range"); return 0; }
	for (int64_t k = 0; k < in.query_aa; ++k)
		if (in.queries[k] >= in.asize) { nsh_set_error("gapless: query residue %lld out of range", (long long)k); return 0; }
	in.n_anch_used = 0;
	for (int32_t i = 0; i < in.n_candidates; ++i) {
		const ns_gpu_gapless_candidate_t &c = in.candidates[i];
		const int64_t na = c.anchor_count < kMaxAnchors? c.anchor_count : kMaxAnchors;
		if (c.target_offset < 0 || c.target_length < 0 || c.target_offset + c.target_length > in.packed_nt_bases ||
				c.query_offset < 0 || c.query_length < 0 || c.query_offset + c.query_length > in.query_aa ||
				(na > 0 && (c.anchor_offset < 0 || c.anchor_offset + na > in.n_anchors))) {
			nsh_set_error("gapless: candidate %d out of range", i);
			return 0;
		}
		in.n_anch_used += na > 0? na : 0;
	}
	return 1;
}

// uploads the batch and queues the scoring kernel on g_stream; the scores stay in g_buf[B_SCORE].
// The caller holds g_mu. Returns the upload's seconds, or a negative value on error.
double queue_scores(const ScoreIn &in)
{
	const size_t sz[B_COUNT] = { (size_t)(in.packed_nt_bases + 1) >> 1, (size_t)in.query_aa,
		(size_t)in.n_candidates * sizeof(*in.candidates), (size_t)in.n_anchors * sizeof(*in.anchors),
		(size_t)in.asize * in.asize, 64, (size_t)in.n_candidates * sizeof(int32_t) };
	for (int b = 0; b < B_COUNT; ++b) if (!ensure(g_buf[b], sz[b])) return -1.0;
#define QSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("gapless: %s: %s", #x, hipGetErrorString(e_)); return -1.0; } } while (0)
	if (!g_stream) QSOK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
	const auto t0 = std::chrono::steady_clock::now();
	if (g_target_src != in.packed_nt || g_target_bases != in.packed_nt_bases) {
		g_target_src = nullptr;
		QSOK(hipMemcpyAsync(g_buf[B_TARGET].p, in.packed_nt, sz[B_TARGET], hipMemcpyHostToDevice, g_stream));
	}
	QSOK(hipMemcpyAsync(g_buf[B_QUERY].p, in.queries, sz[B_QUERY], hipMemcpyHostToDevice, g_stream));
	QSOK(hipMemcpyAsync(g_buf[B_CAND].p, in.candidates, sz[B_CAND], hipMemcpyHostToDevice, g_stream));
	QSOK(hipMemcpyAsync(g_buf[B_ANCH].p, in.anchors, sz[B_ANCH], hipMemcpyHostToDevice, g_stream));
	QSOK(hipMemcpyAsync(g_buf[B_MAT].p, in.score_matrix, sz[B_MAT], hipMemcpyHostToDevice, g_stream));
	QSOK(hipMemcpyAsync(g_buf[B_CODON].p, in.codon, 64, hipMemcpyHostToDevice, g_stream));
	QSOK(hipStreamSynchronize(g_stream));
	g_target_src = in.packed_nt, g_target_bases = in.packed_nt_bases;
	const double up = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	GlArgs a;
	a.pk = (const uint8_t*)g_buf[B_TARGET].p, a.q = (const uint8_t*)g_buf[B_QUERY].p;
	a.cand = (const ns_gpu_gapless_candidate_t*)g_buf[B_CAND].p, a.anch = (const ns_gpu_gapless_anchor_t*)g_buf[B_ANCH].p;
	a.n = in.n_candidates, a.bw = in.block_width, a.kmer = in.kmer, a.xdrop = in.xdrop, a.asize = in.asize;
	a.mat = (const int8_t*)g_buf[B_MAT].p, a.codon = (const uint8_t*)g_buf[B_CODON].p, a.score = (int32_t*)g_buf[B_SCORE].p;
	hipLaunchKernelGGL(gapless_kernel, dim3((in.n_candidates + kWaves - 1) / kWaves), dim3(64 * kWaves), 0, g_stream, a);
	QSOK(hipGetLastError());
#undef QSOK
	nsh_count_launch(1);
	++g_batches, g_cands += in.n_candidates, g_anchors += in.n_anch_used;
	g_seedpos += in.n_anch_used * (uint64_t)in.block_width;
	return up;
}

// the selection inputs of the fused call
enum { S_OFF, S_QLEN, S_CHN, S_VID, S_VS, S_KEEP, S_K, S_COUNT };
DevBuf g_sel[S_COUNT];

} // namespace

extern "C" int ns_gpu_gapless_score_batch(const uint8_t *packed_nt, int64_t packed_nt_bases,
		const uint8_t *queries, int64_t query_aa,
		const ns_gpu_gapless_candidate_t *candidates, int32_t n_candidates,
		const ns_gpu_gapless_anchor_t *anchors, int64_t n_anchors,
		int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *score_matrix, int32_t asize, const uint8_t *codon,
		int32_t *scores)
{
	ScoreIn in = { packed_nt, packed_nt_bases, queries, query_aa, candidates, n_candidates, anchors, n_anchors,
		block_width, kmer, xdrop, score_matrix, asize, codon };
	if (!scores || !check_inputs(in)) return 0;
	std::lock_guard<std::mutex> lk(g_mu);
	const double up = queue_scores(in);
	if (up < 0) return 0;
	auto t1 = std::chrono::steady_clock::now();
	GLOK(hipStreamSynchronize(g_stream));
	auto t2 = std::chrono::steady_clock::now();
	GLOK(hipMemcpyAsync(scores, g_buf[B_SCORE].p, (size_t)n_candidates * 4, hipMemcpyDeviceToHost, g_stream));
	GLOK(hipStreamSynchronize(g_stream));
	auto t3 = std::chrono::steady_clock::now();
	using sec = std::chrono::duration<double>;
	nsh_add_times(up, sec(t2 - t1).count(), sec(t3 - t2).count());
	return 1;
}

extern "C" int nsh_gapless_prefilter_batch(const uint8_t *packed_nt, int64_t packed_nt_bases,
		const uint8_t *queries, int64_t query_aa,
		const ns_gpu_gapless_candidate_t *candidates, int32_t n_candidates,
		const ns_gpu_gapless_anchor_t *anchors, int64_t n_anchors,
		int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *score_matrix, int32_t asize, const uint8_t *codon,
		int32_t n_query, const int32_t *cand_off, const int32_t *qlen, const int32_t *chn_sc, const uint32_t *vid,
		const int64_t *vs, int32_t top_k, int32_t min_per_vid, int32_t adaptive,
		int32_t *scores, uint8_t *keep, int32_t *k_out, nsh_gapless_calib_t *cal)
{
	ScoreIn in = { packed_nt, packed_nt_bases, queries, query_aa, candidates, n_candidates, anchors, n_anchors,
		block_width, kmer, xdrop, score_matrix, asize, codon };
	if (!scores || !keep || n_query < 1 || !cand_off || !qlen || !chn_sc || !vid || !vs || top_k < 1 ||
			min_per_vid < 0 || min_per_vid > top_k) {
		nsh_set_error("gapless prefilter: bad arguments (need top_k >= 1, 0 <= min_per_vid <= top_k)");
		return 0;
	}
	if (cand_off[0] != 0 || cand_off[n_query] != n_candidates) {
		nsh_set_error("gapless prefilter: cand_off must run from 0 to n_candidates");
		return 0;
	}
	int32_t maxn = 0;
	for (int32_t q = 0; q < n_query; ++q) {
		const int32_t n = cand_off[q + 1] - cand_off[q];
		if (n < 0 || n > 8192) { nsh_set_error("gapless prefilter: query %d has %d candidates (0..8192)", q, n); return 0; }
		maxn = std::max(maxn, n);
	}
	if (!check_inputs(in)) return 0;
	std::lock_guard<std::mutex> lk(g_mu);
	if (queue_scores(in) < 0) return 0;
	const size_t nc = (size_t)n_candidates;
	const size_t sz[S_COUNT] = { (size_t)(n_query + 1) * 4, (size_t)n_query * 4, nc * 4, nc * 4, nc * 8, nc,
		(size_t)n_query * 4 };
	GLOK(hipStreamSynchronize(g_stream));   // the selection buffers may be reallocated below
	for (int b = 0; b < S_COUNT; ++b) if (!ensure(g_sel[b], sz[b])) return 0;
	GLOK(hipMemcpyAsync(g_sel[S_OFF].p, cand_off, sz[S_OFF], hipMemcpyHostToDevice, g_stream));
	GLOK(hipMemcpyAsync(g_sel[S_QLEN].p, qlen, sz[S_QLEN], hipMemcpyHostToDevice, g_stream));
	GLOK(hipMemcpyAsync(g_sel[S_CHN].p, chn_sc, sz[S_CHN], hipMemcpyHostToDevice, g_stream));
	GLOK(hipMemcpyAsync(g_sel[S_VID].p, vid, sz[S_VID], hipMemcpyHostToDevice, g_stream));
	GLOK(hipMemcpyAsync(g_sel[S_VS].p, vs, sz[S_VS], hipMemcpyHostToDevice, g_stream));
	if (!nsh_gsel::run(g_stream, n_query, (const int32_t*)g_sel[S_OFF].p, (const int32_t*)g_sel[S_QLEN].p,
			(const int32_t*)g_buf[B_SCORE].p, (const int32_t*)g_sel[S_CHN].p, (const uint32_t*)g_sel[S_VID].p,
			(const int64_t*)g_sel[S_VS].p, n_candidates, maxn, top_k, min_per_vid, adaptive,
			(uint8_t*)g_sel[S_KEEP].p, (int32_t*)g_sel[S_K].p, cal))
		return 0;
	GLOK(hipMemcpyAsync(scores, g_buf[B_SCORE].p, nc * 4, hipMemcpyDeviceToHost, g_stream));
	GLOK(hipMemcpyAsync(keep, g_sel[S_KEEP].p, nc, hipMemcpyDeviceToHost, g_stream));
	if (k_out) GLOK(hipMemcpyAsync(k_out, g_sel[S_K].p, sz[S_K], hipMemcpyDeviceToHost, g_stream));
	GLOK(hipStreamSynchronize(g_stream));
	return 1;
}
