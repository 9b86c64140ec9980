// The profile prescreen search (prescreen.c:133-486) on the device: an ORF database and spaced
// k-mer index of a genome, then per profile the similar-k-mer hits, the double-hit diagonals,
// their ungapped scores, the best diagonal per ORF, the banded gapped score and the E-value.
// The candidates equal the reference worker's (ps_query_worker) in content and order.
//
// Index (mp_prescreen_build): the six-frame ORFs come from the ORF walk kernels (orf_scan.hip),
// strand by strand in the reference's order (contig, then strand). Every window of `span`
// residues of an ORF gets a spaced k-mer code (the mmseqs patterns). The windows are counted per
// code with atomics, the counts are turned into bucket starts by a device scan, and the
// windows are scattered (orf << 32 | position). The order inside a bucket does not matter,
// because the hits are sorted before use.
//
// Search (ps_query_worker, for a batch of profiles at once as mp_prescreen_batch runs it):
//   * each profile row's 20 residues are ranked by score, stably, as the reference's insertion
//     sort does;
//   * the k-mers scoring at least thr of every query window of every profile are enumerated level
//     by level (one pattern position per level) instead of one depth-first search per window:
//     each level keeps the prefixes that can still reach thr, in the depth-first order, and drops
//     those whose rank inside their window is past the cap (kmer_max), so the k-mers kept per
//     window are exactly the reference's first kmer_max. Then every kept k-mer's index bucket is
//     counted, the counts scanned, and the hits written as sort keys
//     (profile | ORF | diagonal + max query length);
//   * a device radix sort orders the keys (the reference's order within each profile);
//   * runs of at least two equal keys are the double-hit diagonals, compacted in order;
//   * one lane per diagonal computes ps_ungapped;
//   * per (profile, ORF), the first strictly best diagonal is kept when it reaches min_ungapped.
// Each profile's short candidate list is then truncated to max_seqs by (score desc, ORF asc) when
// it is longer. The banded gapped scores come from ps_gapped.hip's kernel on the resident
// residues, all profiles in one launch per column count. The E-value is a few double operations
// per candidate with the C library's exp/erfc; it is evaluated on the host so that its bits
// equal the reference's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int PS_AA = 20;
constexpr uint32_t DIAG_OFF = 1u << 30;

// mmseqs spaced seeds (the reference's ps_pat5/6/7): 1 marks an informative position
const int8_t kPat5[] = {1,1,0,0,1,0,0,0,0,1,0,1};
const int8_t kPat6[] = {1,1,0,1,0,1,0,0,1,1};
const int8_t kPat7[] = {1,1,0,1,0,1,1,0,0,1,1};

struct Pat { int32_t kmer, span; int8_t pos[8]; };

__device__ __forceinline__ uint32_t kcode(const uint8_t *s, const Pat &p)
{
	uint32_t c = 0, mul = 1;
	for (int t = 0; t < p.kmer; ++t, mul *= PS_AA) c += s[p.pos[t]] * mul;
	return c;
}

// one wave per ORF: count (PASS 0) or place (PASS 1) every window's k-mer
template <int PASS>
__global__ void __launch_bounds__(256) index_kernel(const uint8_t *aa, const int64_t *orf_off, const int32_t *orf_len,
		int64_t n_orf, Pat p, unsigned long long *cnt, const int64_t *off, uint64_t *ent)
{
	const int64_t o = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
	if (o >= n_orf) return;
	const uint8_t *s = aa + orf_off[o];
	const int32_t n = orf_len[o];
	for (int32_t q = threadIdx.x & 63; q + p.span <= n; q += 64) {
		const uint32_t c = kcode(s + q, p);
		if (PASS == 0) atomicAdd(&cnt[c], 1ull);
		else ent[off[c] + (int64_t)atomicAdd(&cnt[c], 1ull)] = (uint64_t)o << 32 | (uint32_t)q;
	}
}

// per profile row: the 20 residues by decreasing score, ties in residue order (the reference's
// stable insertion sort); one thread per (row, residue) places its residue at its rank
__global__ void rank_kernel(const int8_t *pssm, int64_t rows, int8_t *sc, uint8_t *ai)
{
	const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= rows * PS_AA) return;
	const int64_t i = g / PS_AA;
	const int a = (int)(g - i * PS_AA);
	const int8_t *row = pssm + i * 22;
	const int8_t x = row[a];
	int rank = 0;
#pragma unroll
	for (int b = 0; b < PS_AA; ++b) {
		const int8_t y = row[b];
		rank += y > x || (y == x && b < a);
	}
	sc[i * PS_AA + rank] = x, ai[i * PS_AA + rank] = (uint8_t)a;
}

// The similar k-mers of every query window, level by level. A prefix fixes the ranks chosen at the
// first t pattern positions; the level-t frontier lists the prefixes that can still reach thr,
// ordered by window, then by their ranks (the reference's depth-first order). Since each ranked
// row decreases, the ranks a prefix can take at position t form a leading run of the row, found
// by scanning until the bound fails, exactly where the depth-first search backtracks. Every
// prefix on the frontier has at least one completion, so a prefix whose index inside its window is
// >= cap has all of its completions past the cap: dropping those keeps each window's frontier
// within cap entries, and at the last level it keeps exactly the first cap k-mers.
struct LvlArgs {
	const int8_t *sc; const uint8_t *ai;
	const int64_t *wrow;               // per window: PSSM row of its first pattern position
	const int32_t *wq, *wprof;         // per window: query offset and profile
	Pat p; int32_t t, thr;
	int64_t n, cap;
	uint32_t mul;                      // 20^t
	const int32_t *win; const uint32_t *code; const int32_t *sum;   // level-t frontier
	int64_t *cnt; const int64_t *E; int64_t *wstart, *hc;
	int32_t *nwin; uint32_t *ncode; int32_t *nsum;                  // level-(t+1) frontier
	const int64_t *off; const uint64_t *ent;                        // the k-mer index
	uint64_t *hit; int32_t D, OD; int64_t dbias;                    // hit keys: prof << OD | orf << D | diag + dbias
};

__global__ void lvl_init_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	a.nwin[i] = (int32_t)i, a.ncode[i] = 0, a.nsum[i] = 0;
}

// cnt[i]: how many ranks at position t keep prefix i able to reach thr; cnt[n] = 0 closes the scan
__global__ void lvl_count_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i > a.n) return;
	if (i == a.n) { a.cnt[i] = 0; return; }
	const int64_t r0 = a.wrow[a.win[i]];
	int rest = 0, pt = 0;
#pragma unroll
	for (int u = 0; u < 8; ++u) {
		if (u > a.t && u < a.p.kmer) rest += a.sc[(r0 + a.p.pos[u]) * PS_AA];
		if (u == a.t) pt = a.p.pos[u];
	}
	const int8_t *row = a.sc + (r0 + pt) * PS_AA;
	const int s = a.sum[i] + rest;
	int c = 0;
	while (c < PS_AA && s + row[c] >= a.thr) ++c;
	a.cnt[i] = c;
}

// the scanned offset of each window's first prefix
__global__ void lvl_wstart_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	if (i == 0 || a.win[i - 1] != a.win[i]) a.wstart[a.win[i]] = a.E[i];
}

// keep the extensions whose index inside their window is below cap
__global__ void lvl_clamp_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	const int64_t room = a.cap - (a.E[i] - a.wstart[a.win[i]]), c = a.cnt[i];
	a.cnt[i] = room <= 0? 0 : c < room? c : room;
}

__device__ __forceinline__ int64_t pos_row(const LvlArgs &a, int64_t r0)
{
	int pt = 0;
#pragma unroll
	for (int u = 0; u < 8; ++u) if (u == a.t) pt = a.p.pos[u];
	return (r0 + pt) * PS_AA;
}

// the level-(t+1) frontier, at the scanned offsets E
__global__ void lvl_emit_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	const int c = (int)a.cnt[i];
	if (c == 0) return;
	const int32_t w = a.win[i], s = a.sum[i];
	const uint32_t cd = a.code[i];
	const int64_t r = pos_row(a, a.wrow[w]), o = a.E[i];
	for (int j = 0; j < c; ++j)
		a.nwin[o + j] = w, a.ncode[o + j] = cd + a.ai[r + j] * a.mul, a.nsum[o + j] = s + a.sc[r + j];
}

// last level: index entries of the kept k-mers of prefix i (cnt[i] leading ranks); cnt[n] = 0
__global__ void leaf_count_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i > a.n) return;
	if (i == a.n) { a.hc[i] = 0; return; }
	const int c = (int)a.cnt[i];
	int64_t h = 0;
	if (c) {
		const int64_t r = pos_row(a, a.wrow[a.win[i]]);
		const uint32_t cd = a.code[i];
		for (int j = 0; j < c; ++j) {
			const uint32_t k = cd + a.ai[r + j] * a.mul;
			h += a.off[k + 1] - a.off[k];
		}
	}
	a.hc[i] = h;
}

// last level: the hits of prefix i at the scanned offset E[i], as sort keys
__global__ void leaf_emit_kernel(LvlArgs a)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	const int c = (int)a.cnt[i];
	if (c == 0) return;
	const int32_t w = a.win[i];
	const int64_t r = pos_row(a, a.wrow[w]), bias = a.dbias - a.wq[w];
	const uint64_t hi = (uint64_t)a.wprof[w] << a.OD;
	const uint32_t cd = a.code[i];
	int64_t o = a.E[i];
	for (int j = 0; j < c; ++j) {
		const uint32_t k = cd + a.ai[r + j] * a.mul;
		const int64_t b1 = a.off[k + 1];
		for (int64_t e = a.off[k]; e < b1; ++e) {
			const uint64_t x = a.ent[e];
			a.hit[o++] = hi | (x >> 32) << a.D | (uint64_t)((int64_t)(uint32_t)x + bias);
		}
	}
}

// first element of each run of >= 2 equal hits: a double-hit diagonal
__global__ void dbl_flag_kernel(const uint64_t *h, int64_t n, uint8_t *f)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	f[k] = k + 1 < n && h[k] == h[k + 1] && (k == 0 || h[k - 1] != h[k]);
}

struct UngArgs {
	const int8_t *pssm; const int64_t *prow; const int32_t *plen;
	const uint8_t *aa; const int64_t *orf_off; const int32_t *orf_len;
	const uint64_t *dg; int64_t n;
	int32_t D, OD; int64_t dbias;
	int32_t *ung, *qs, *qe;
};

// ps_ungapped (prescreen.c:334-344) of each double-hit diagonal
__global__ void ung_kernel(UngArgs a)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= a.n) return;
	const uint64_t key = a.dg[k];
	const int32_t pf = (int32_t)(key >> a.OD);
	const int64_t o = (int64_t)((key >> a.D) & ((1ull << (a.OD - a.D)) - 1));
	const int32_t diag = (int32_t)((int64_t)(key & ((1ull << a.D) - 1)) - a.dbias), n = a.orf_len[o], L = a.plen[pf];
	const int8_t *pssm = a.pssm + a.prow[pf] * 22;
	const uint8_t *orf = a.aa + a.orf_off[o];
	const int32_t i0 = diag < 0? -diag : 0, i1 = n - diag < L? n - diag : L;
	int32_t sc = 0, best = 0, st = i0, qs = i0, qe = i0, i = i0;
	for (; i + 8 <= i1; i += 8) {   // eight residues, then their eight scores, in flight together
		int r[8], v[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) r[u] = orf[i + u + diag];
#pragma unroll
		for (int u = 0; u < 8; ++u) v[u] = pssm[(int64_t)(i + u) * 22 + r[u]];
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			sc += v[u];
			if (sc < 0) sc = 0, st = i + u + 1;
			else if (sc > best) best = sc, qs = st, qe = i + u + 1;
		}
	}
	for (; i < i1; ++i) {
		sc += pssm[(int64_t)i * 22 + orf[i + diag]];
		if (sc < 0) sc = 0, st = i + 1;
		else if (sc > best) best = sc, qs = st, qe = i + 1;
	}
	a.ung[k] = best, a.qs[k] = qs, a.qe[k] = qe;
}

// per (profile, ORF) group (consecutive, diagonals increasing): the first strictly best
// diagonal; keep[k] marks the group's first slot, which receives the winner, when that best
// reaches min_ungapped
__global__ void best_kernel(uint64_t *dg, int64_t n, int32_t D, int32_t *ung, int32_t *qs, int32_t *qe,
		int32_t min_ungapped, uint8_t *keep)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const uint64_t g = dg[k] >> D;
	if (k > 0 && (dg[k - 1] >> D) == g) { keep[k] = 0; return; }
	int32_t best = -1;
	int64_t bj = k;
	for (int64_t j = k; j < n && (dg[j] >> D) == g; ++j)
		if (ung[j] > best) best = ung[j], bj = j;
	keep[k] = best >= min_ungapped;
	if (bj != k) {   // only this thread touches the group
		dg[k] = dg[bj];
		ung[k] = ung[bj], qs[k] = qs[bj], qe[k] = qe[bj];
	}
}

struct Kept { uint64_t key; int32_t ung, qs, qe, pad; };

__global__ void gather_kernel(const int32_t *sel, int64_t n, const uint64_t *dg, const int32_t *ung, const int32_t *qs,
		const int32_t *qe, Kept *out)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int32_t j = sel[k];
	out[k].key = dg[j], out[k].ung = ung[j], out[k].qs = qs[j], out[k].qe = qe[j], out[k].pad = 0;
}

// mmseqs2 / ALP statistics for BLOSUM62 with gap costs 11/1 in half-bits (the values the
// reference states at prescreen.c:225-233): lambda, K, and the finite-size correction terms
constexpr double kLambda = 0.27359865037097330642, kK = 0.044620920658722244834;
constexpr double kA = 1.5938724404943873658, kB = -19.959867650284412122;
constexpr double kAlpha = 30.455610143099914211, kBeta = -622.28684628915891608;
constexpr double kSigma = 29.602444874818868215, kTau = -601.81087985041381216;

// Karlin-Altschul E-value with the ALP finite-size correction of the search area
// (get_appr_tail_prob_with_cov_without_errors): query length m, database residues n
double evalue(double S, double m, double n)
{
	const double inv_sqrt_2pi = 0.39894228040143267794;
	const double v_min = 2.0 * kAlpha / kLambda, c_min = 2.0 * kSigma / kLambda;
	const double l = kA * S + kB;
	double v = kAlpha * S + kBeta;
	if (v < v_min) v = v_min;
	const double sv = std::sqrt(v), mF = (m - l) / sv, nF = (n - l) / sv;
	const double Pm = 0.5 * std::erfc(-mF * 0.70710678118654752440), Pn = 0.5 * std::erfc(-nF * 0.70710678118654752440);
	const double p1 = (m - l) * Pm + sv * inv_sqrt_2pi * std::exp(-0.5 * mF * mF);
	const double p2 = (n - l) * Pn + sv * inv_sqrt_2pi * std::exp(-0.5 * nF * nF);
	double cy = kSigma * S + kTau;
	if (cy < c_min) cy = c_min;
	double area = p1 * p2 + cy * Pm * Pn;
	if (area < 1e-300) area = 1e-300;
	return kK * area * std::exp(-kLambda * S);
}

} // namespace

struct nsh_prescreen_s {
	Pat pat;
	int64_t n_orf = 0, n_aa = 0, n_bucket = 0, n_ent = 0;
	std::vector<uint32_t> vid;
	std::vector<int64_t> vs, aa_off;
	std::vector<int32_t> aa_len;
	uint8_t *d_aa = nullptr;
	int64_t *d_orf_off = nullptr, *d_off = nullptr;
	int32_t *d_orf_len = nullptr;
	uint64_t *d_ent = nullptr;
	int32_t max_orf_len = 0;
	hipStream_t stream = nullptr;
	std::mutex mu;
	void *ws[32] = {};          // search workspace, grown on demand and kept
	size_t ws_bytes[32] = {};
	void *pin[4] = {};          // pinned host staging of the search's transfers
	size_t pin_bytes[4] = {};
	int64_t stats[7] = {};      // last search: windows, k-mers, index hits, double-hit diagonals,
	                            // ORFs >= min_ungapped, aligned, kept
};

#define PSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("prescreen: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" void nsh_prescreen_destroy(nsh_prescreen_t *ps)
{
	if (!ps) return;
	if (ps->stream) (void)hipStreamSynchronize(ps->stream);
	void *bufs[] = { ps->d_aa, ps->d_orf_off, ps->d_off, ps->d_orf_len, ps->d_ent };
	for (void *b : bufs) if (b) (void)hipFree(b);
	for (void *b : ps->ws) if (b) (void)hipFree(b);
	for (void *b : ps->pin) if (b) (void)hipHostFree(b);
	if (ps->stream) (void)hipStreamDestroy(ps->stream);
	delete ps;
}

extern "C" nsh_prescreen_t *nsh_prescreen_build(int32_t n_ctg, const uint8_t *nt, const int64_t *ctg_off,
		const int64_t *ctg_len, int32_t kmer, int32_t min_orf)
{
	if (!nsh_ensure_device()) return nullptr;
	if (n_ctg < 0 || (n_ctg > 0 && (!nt || !ctg_off || !ctg_len))) { nsh_set_error("prescreen: bad arguments"); return nullptr; }
	nsh_prescreen_t *ps = new nsh_prescreen_s();
	const int8_t *pat;
	int plen;
	kmer = kmer < 5? 5 : kmer > 7? 7 : kmer;
	if (kmer == 5) pat = kPat5, plen = (int)sizeof(kPat5);
	else if (kmer == 7) pat = kPat7, plen = (int)sizeof(kPat7);
	else pat = kPat6, plen = (int)sizeof(kPat6);
	ps->pat.kmer = kmer, ps->pat.span = plen;
	for (int i = 0, c = 0; i < plen; ++i) if (pat[i]) ps->pat.pos[c++] = (int8_t)i;
	const int32_t morf = min_orf > plen? min_orf : plen;
	std::vector<uint8_t> aa;
	unsigned long long *d_cnt = nullptr;
	void *d_tmp = nullptr;
	size_t tmp_bytes = 0;
	// ORFs, strand by strand (vid = contig * 2 + reverse), as ps_orf_worker produces them
	for (int32_t c = 0; c < n_ctg; ++c) {
		if (ctg_off[c] < 0 || ctg_len[c] < 0) { nsh_set_error("prescreen: contig %d out of range", c); delete ps; return nullptr; }
		std::vector<uint8_t> s(nt + ctg_off[c], nt + ctg_off[c] + ctg_len[c]);
		for (int rev = 0; rev < 2; ++rev) {
			if (rev) {
				std::reverse(s.begin(), s.end());
				for (auto &b : s) b = b >= 4? b : (uint8_t)(3 - b);
			}
			const int64_t cap = (int64_t)s.size() + 8;
			std::vector<int64_t> st(cap);
			std::vector<int32_t> len(cap);
			std::vector<uint8_t> res(cap);
			int64_t tot = 0;
			const int64_t n = nsh_orf_scan(s.data(), (int64_t)s.size(), morf, cap, st.data(), len.data(), cap, res.data(), &tot);
			if (n < 0) { delete ps; return nullptr; }
			for (int64_t k = 0, o = 0; k < n; ++k) {
				ps->vid.push_back((uint32_t)(c * 2 + rev)), ps->vs.push_back(st[k]), ps->aa_len.push_back(len[k]);
				ps->aa_off.push_back((int64_t)aa.size() + o);
				o += len[k];
			}
			aa.insert(aa.end(), res.begin(), res.begin() + tot);
		}
	}
	ps->n_orf = (int64_t)ps->vid.size(), ps->n_aa = (int64_t)aa.size();
	for (const int32_t l : ps->aa_len) ps->max_orf_len = std::max(ps->max_orf_len, l);
	ps->n_bucket = 1;
	for (int i = 0; i < kmer; ++i) ps->n_bucket *= PS_AA;
	{
		PSOK(hipStreamCreateWithFlags(&ps->stream, hipStreamNonBlocking));
		PSOK(hipMalloc((void**)&ps->d_aa, (size_t)ps->n_aa + 8));
		PSOK(hipMalloc((void**)&ps->d_orf_off, (size_t)ps->n_orf * 8 + 8));
		PSOK(hipMalloc((void**)&ps->d_orf_len, (size_t)ps->n_orf * 4 + 8));
		PSOK(hipMalloc((void**)&ps->d_off, (size_t)(ps->n_bucket + 1) * 8));
		PSOK(hipMalloc((void**)&d_cnt, (size_t)(ps->n_bucket + 1) * 8));
		if (ps->n_aa) PSOK(hipMemcpyAsync(ps->d_aa, aa.data(), (size_t)ps->n_aa, hipMemcpyHostToDevice, ps->stream));
		if (ps->n_orf) {
			PSOK(hipMemcpyAsync(ps->d_orf_off, ps->aa_off.data(), (size_t)ps->n_orf * 8, hipMemcpyHostToDevice, ps->stream));
			PSOK(hipMemcpyAsync(ps->d_orf_len, ps->aa_len.data(), (size_t)ps->n_orf * 4, hipMemcpyHostToDevice, ps->stream));
		}
		PSOK(hipMemsetAsync(d_cnt, 0, (size_t)(ps->n_bucket + 1) * 8, ps->stream));
		const unsigned grid = (unsigned)((ps->n_orf + 3) / 4);
		if (grid)
			hipLaunchKernelGGL(index_kernel<0>, dim3(grid), dim3(256), 0, ps->stream, ps->d_aa, ps->d_orf_off, ps->d_orf_len,
					ps->n_orf, ps->pat, d_cnt, (const int64_t*)nullptr, (uint64_t*)nullptr);
		PSOK(hipGetLastError());
		PSOK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (const int64_t*)d_cnt, ps->d_off, (int)(ps->n_bucket + 1), ps->stream));
		PSOK(hipMalloc(&d_tmp, tmp_bytes));
		PSOK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, (const int64_t*)d_cnt, ps->d_off, (int)(ps->n_bucket + 1), ps->stream));
		PSOK(hipMemcpyAsync(&ps->n_ent, ps->d_off + ps->n_bucket, 8, hipMemcpyDeviceToHost, ps->stream));
		PSOK(hipStreamSynchronize(ps->stream));
		PSOK(hipMalloc((void**)&ps->d_ent, (size_t)ps->n_ent * 8 + 8));
		PSOK(hipMemsetAsync(d_cnt, 0, (size_t)(ps->n_bucket + 1) * 8, ps->stream));
		if (grid)
			hipLaunchKernelGGL(index_kernel<1>, dim3(grid), dim3(256), 0, ps->stream, ps->d_aa, ps->d_orf_off, ps->d_orf_len,
					ps->n_orf, ps->pat, d_cnt, (const int64_t*)ps->d_off, ps->d_ent);
		PSOK(hipGetLastError());
		PSOK(hipStreamSynchronize(ps->stream));
		nsh_count_launch(3);
	}
	(void)hipFree(d_cnt), (void)hipFree(d_tmp);
	return ps;
fail:
	if (d_cnt) (void)hipFree(d_cnt);
	if (d_tmp) (void)hipFree(d_tmp);
	nsh_prescreen_destroy(ps);
	return nullptr;
}

extern "C" int64_t nsh_prescreen_size(const nsh_prescreen_t *ps, int32_t what)
{
	if (!ps) return -1;
	return what == 0? ps->n_orf : what == 1? ps->n_aa : ps->n_ent;
}

namespace {

enum { W_IN, W_SC, W_AI, W_F0W, W_F0C, W_F0S, W_F1W, W_F1C, W_F1S, W_CNT, W_E, W_WSTART, W_HC, W_HO, W_HIT, W_HIT2,
	W_FLAG, W_DG, W_NSEL, W_UNG, W_QS, W_QE, W_SEL, W_KEPT, W_TMP, W_GAP, W_N };
static_assert(W_N <= 32, "workspace slots");

constexpr int64_t kMaxItems = int64_t(1) << 28;   // frontier entries or hits of one chunk

// a workspace buffer of at least count T, kept by the index between calls
template <class T>
hipError_t wsget(nsh_prescreen_t *ps, int slot, int64_t count, T **out)
{
	const size_t bytes = (size_t)(count < 1? 1 : count) * sizeof(T) + 64;
	if (ps->ws_bytes[slot] < bytes) {
		hipError_t e = hipStreamSynchronize(ps->stream);
		if (e != hipSuccess) return e;
		if (ps->ws[slot]) (void)hipFree(ps->ws[slot]);
		ps->ws[slot] = nullptr, ps->ws_bytes[slot] = 0;
		const size_t grow = std::max(bytes, ps->ws_bytes[slot] + ps->ws_bytes[slot] / 2);
		e = hipMalloc(&ps->ws[slot], grow);
		if (e != hipSuccess) return e;
		ps->ws_bytes[slot] = grow;
	}
	*out = (T*)ps->ws[slot];
	return hipSuccess;
}

enum { P_IN, P_SCALAR, P_OUT, P_N };

// pinned host memory of at least bytes, kept by the index between calls
hipError_t pinget(nsh_prescreen_t *ps, int slot, size_t bytes, void **out)
{
	bytes += 64;
	if (ps->pin_bytes[slot] < bytes) {
		hipError_t e = hipStreamSynchronize(ps->stream);
		if (e != hipSuccess) return e;
		if (ps->pin[slot]) (void)hipHostFree(ps->pin[slot]);
		ps->pin[slot] = nullptr, ps->pin_bytes[slot] = 0;
		const size_t grow = std::max(bytes, (size_t)4096);
		e = hipHostMalloc(&ps->pin[slot], grow, hipHostMallocDefault);
		if (e != hipSuccess) return e;
		ps->pin_bytes[slot] = grow;
	}
	*out = ps->pin[slot];
	return hipSuccess;
}

// the stream's value of a device int64 (through pinned memory; waits for the stream)
hipError_t read_i64(nsh_prescreen_t *ps, const int64_t *d, int64_t *v)
{
	void *h;
	hipError_t e = pinget(ps, P_SCALAR, 8, &h);
	if (e != hipSuccess) return e;
	e = hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, ps->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(ps->stream);
	if (e == hipSuccess) *v = *(const int64_t*)h;
	return e;
}

size_t al64(size_t x) { return (x + 63) & ~(size_t)63; }

int bits_for(uint64_t x) { return x? 64 - __builtin_clzll(x) : 0; }

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

struct SearchIn {
	const int8_t *pssm; const int32_t *L; const int64_t *prow;   // host: profiles and their first rows
	int32_t thr, kmer_max, min_ungapped, max_seqs, band, go, ge;
	double max_evalue;
};

#define WSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("prescreen: %s: %s", #x, hipGetErrorString(e_)); return 0; } } while (0)

// the candidates of profiles [p0, p1) into res[p - p0]: 1 done, 2 too large for one chunk, 0 error
int search_chunk(nsh_prescreen_t *ps, const SearchIn &in, int32_t p0, int32_t p1,
		std::vector<std::vector<nsh_ps_cand_t>> &res)
{
	hipStream_t s = ps->stream;
	const int32_t np = p1 - p0, span = ps->pat.span, k = ps->pat.kmer;
	const int64_t rows = in.prow[p1] - in.prow[p0];
	int32_t maxL = 0;
	std::vector<int64_t> prow(np);
	std::vector<int32_t> plen(np);
	std::vector<int64_t> wrow;
	std::vector<int32_t> wq, wprof;
	for (int32_t p = 0; p < np; ++p) {
		prow[p] = in.prow[p0 + p] - in.prow[p0], plen[p] = in.L[p0 + p];
		maxL = std::max(maxL, plen[p]);
		for (int32_t q = 0; q + span <= plen[p]; ++q) wrow.push_back(prow[p] + q), wq.push_back(q), wprof.push_back(p);
	}
	const int64_t W = (int64_t)wrow.size();
	if (W == 0) return 1;
	if (W > kMaxItems) return 2;
	int64_t st[7] = { W, 0, 0, 0, 0, 0, 0 };   // added to the index's statistics when the chunk completes
	struct Commit {
		nsh_prescreen_t *ps; int64_t *st; bool on = true;
		~Commit() { if (on) for (int i = 0; i < 7; ++i) ps->stats[i] += st[i]; }
	} commit = { ps, st };
	// sort-key layout: profile | ORF | diagonal + maxL (the diagonal lies in (-maxL, max_orf_len))
	const int D = std::max(1, bits_for((uint64_t)maxL + (uint64_t)ps->max_orf_len));
	const int O = std::max(1, bits_for((uint64_t)(ps->n_orf - 1)));
	const int P = bits_for((uint64_t)(np - 1));
	if (P + O + D > 63) { commit.on = false; return 2; }   // retried at half size: no counts from this attempt
	int8_t *d_pssm, *d_sc;
	uint8_t *d_ai, *d_in;
	int64_t *d_prow, *d_wrow;
	int32_t *d_plen, *d_wq, *d_wprof;
	{   // the inputs in one pinned block and one copy: PSSM rows | prow | wrow | plen | wq | wprof
		const size_t o_prow = al64((size_t)rows * 22), o_wrow = o_prow + al64((size_t)np * 8);
		const size_t o_plen = o_wrow + al64((size_t)W * 8), o_wq = o_plen + al64((size_t)np * 4);
		const size_t o_wprof = o_wq + al64((size_t)W * 4), tot = o_wprof + al64((size_t)W * 4);
		void *h;
		WSOK(pinget(ps, P_IN, tot, &h));
		uint8_t *hb = (uint8_t*)h;
		std::memcpy(hb, in.pssm + in.prow[p0] * 22, (size_t)rows * 22);
		std::memcpy(hb + o_prow, prow.data(), (size_t)np * 8);
		std::memcpy(hb + o_wrow, wrow.data(), (size_t)W * 8);
		std::memcpy(hb + o_plen, plen.data(), (size_t)np * 4);
		std::memcpy(hb + o_wq, wq.data(), (size_t)W * 4);
		std::memcpy(hb + o_wprof, wprof.data(), (size_t)W * 4);
		WSOK(wsget(ps, W_IN, (int64_t)tot, &d_in));
		WSOK(hipMemcpyAsync(d_in, hb, tot, hipMemcpyHostToDevice, s));
		d_pssm = (int8_t*)d_in, d_prow = (int64_t*)(d_in + o_prow), d_wrow = (int64_t*)(d_in + o_wrow);
		d_plen = (int32_t*)(d_in + o_plen), d_wq = (int32_t*)(d_in + o_wq), d_wprof = (int32_t*)(d_in + o_wprof);
	}
	WSOK(wsget(ps, W_SC, rows * PS_AA, &d_sc));
	WSOK(wsget(ps, W_AI, rows * PS_AA, &d_ai));
	hipLaunchKernelGGL(rank_kernel, dim3(blocks(rows * PS_AA)), dim3(256), 0, s, d_pssm, rows, d_sc, d_ai);
	WSOK(hipGetLastError());

	LvlArgs la = {};
	la.sc = d_sc, la.ai = d_ai, la.wrow = d_wrow, la.wq = d_wq, la.wprof = d_wprof, la.p = ps->pat;
	la.thr = in.thr, la.cap = in.kmer_max, la.off = ps->d_off, la.ent = ps->d_ent;
	la.D = D, la.OD = O + D, la.dbias = maxL;
	int cur = 0;
	const int fw[2] = { W_F0W, W_F1W }, fc[2] = { W_F0C, W_F1C }, fs[2] = { W_F0S, W_F1S };
	int32_t *win; uint32_t *code; int32_t *sum;
	WSOK(wsget(ps, fw[0], W, &win));
	WSOK(wsget(ps, fc[0], W, &code));
	WSOK(wsget(ps, fs[0], W, &sum));
	la.n = W, la.nwin = win, la.ncode = code, la.nsum = sum;
	hipLaunchKernelGGL(lvl_init_kernel, dim3(blocks(W)), dim3(256), 0, s, la);
	int64_t *d_wstart, *d_cnt, *d_E;
	WSOK(wsget(ps, W_WSTART, W, &d_wstart));
	void *d_tmp = nullptr;
	size_t tmp_bytes = 0;
	int64_t n = W, n_hit = 0;
	uint32_t mul = 1;
	for (int t = 0; t < k; ++t, mul *= PS_AA) {
		WSOK(wsget(ps, W_CNT, n + 1, &d_cnt));
		WSOK(wsget(ps, W_E, n + 1, &d_E));
		WSOK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt, d_E, (int)(n + 1), s));
		WSOK(wsget(ps, W_TMP, (int64_t)tmp_bytes, (uint8_t**)&d_tmp));
		la.t = t, la.n = n, la.mul = mul, la.win = win, la.code = code, la.sum = sum;
		la.cnt = d_cnt, la.E = d_E, la.wstart = d_wstart;
		hipLaunchKernelGGL(lvl_count_kernel, dim3(blocks(n + 1)), dim3(256), 0, s, la);
		WSOK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_cnt, d_E, (int)(n + 1), s));
		hipLaunchKernelGGL(lvl_wstart_kernel, dim3(blocks(n)), dim3(256), 0, s, la);
		hipLaunchKernelGGL(lvl_clamp_kernel, dim3(blocks(n)), dim3(256), 0, s, la);
		WSOK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_cnt, d_E, (int)(n + 1), s));
		WSOK(hipGetLastError());
		int64_t tot = 0;
		WSOK(read_i64(ps, d_E + n, &tot));
		nsh_count_launch(4);
		if (t + 1 == k) st[1] = tot;
		if (tot == 0) return 1;
		if (t + 1 < k) {   // the next frontier
			if (tot > kMaxItems) { commit.on = false; return 2; }
			const int nx = cur ^ 1;
			WSOK(wsget(ps, fw[nx], tot, &la.nwin));
			WSOK(wsget(ps, fc[nx], tot, &la.ncode));
			WSOK(wsget(ps, fs[nx], tot, &la.nsum));
			hipLaunchKernelGGL(lvl_emit_kernel, dim3(blocks(n)), dim3(256), 0, s, la);
			WSOK(hipGetLastError());
			nsh_count_launch(1);
			win = la.nwin, code = la.ncode, sum = la.nsum, cur = nx, n = tot;
			continue;
		}
		// the kept k-mers: hit counts, offsets, keys
		int64_t *d_hc;
		WSOK(wsget(ps, W_HC, n + 1, &d_hc));
		la.hc = d_hc;
		hipLaunchKernelGGL(leaf_count_kernel, dim3(blocks(n + 1)), dim3(256), 0, s, la);
		// d_cnt (the kept ranks) stays; the hit offsets go to a second scan buffer
		int64_t *d_ho;
		WSOK(wsget(ps, W_HO, n + 1, &d_ho));
		WSOK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_hc, d_ho, (int)(n + 1), s));
		WSOK(read_i64(ps, d_ho + n, &n_hit));
		nsh_count_launch(2);
		st[2] = n_hit;
		if (n_hit == 0) return 1;
		if (n_hit > kMaxItems) { commit.on = false; return 2; }
		uint64_t *d_hit;
		WSOK(wsget(ps, W_HIT, n_hit, &d_hit));
		la.E = d_ho, la.hit = d_hit;
		hipLaunchKernelGGL(leaf_emit_kernel, dim3(blocks(n)), dim3(256), 0, s, la);
		WSOK(hipGetLastError());
		nsh_count_launch(1);
	}
	// sort, then the double-hit diagonals in order
	uint64_t *d_hit = la.hit, *d_hit2, *d_dg;
	uint8_t *d_flag;
	int64_t *d_nsel, n_dg = 0;
	const int end_bit = P + O + D;
	WSOK(wsget(ps, W_HIT2, n_hit, &d_hit2));
	WSOK(wsget(ps, W_FLAG, n_hit, &d_flag));
	WSOK(wsget(ps, W_DG, n_hit, &d_dg));
	WSOK(wsget(ps, W_NSEL, 1, &d_nsel));
	size_t need = 0;
	WSOK(hipcub::DeviceRadixSort::SortKeys(nullptr, need, d_hit, d_hit2, (int)n_hit, 0, end_bit, s));
	tmp_bytes = need;
	WSOK(hipcub::DeviceSelect::Flagged(nullptr, need, d_hit2, d_flag, d_dg, d_nsel, (int)n_hit, s));
	tmp_bytes = std::max(tmp_bytes, need);
	WSOK(hipcub::DeviceSelect::Flagged(nullptr, need, hipcub::CountingInputIterator<int32_t>(0), d_flag,
			(int32_t*)nullptr, d_nsel, (int)n_hit, s));
	tmp_bytes = std::max(tmp_bytes, need);
	WSOK(wsget(ps, W_TMP, (int64_t)tmp_bytes, (uint8_t**)&d_tmp));
	WSOK(hipcub::DeviceRadixSort::SortKeys(d_tmp, tmp_bytes, d_hit, d_hit2, (int)n_hit, 0, end_bit, s));
	hipLaunchKernelGGL(dbl_flag_kernel, dim3(blocks(n_hit)), dim3(256), 0, s, d_hit2, n_hit, d_flag);
	WSOK(hipcub::DeviceSelect::Flagged(d_tmp, tmp_bytes, d_hit2, d_flag, d_dg, d_nsel, (int)n_hit, s));
	WSOK(read_i64(ps, d_nsel, &n_dg));
	nsh_count_launch(3);
	st[3] = n_dg;
	if (n_dg == 0) return 1;
	// ungapped scores, the best diagonal per (profile, ORF), the kept ones packed for the host
	int32_t *d_ung, *d_qs, *d_qe, *d_sel;
	Kept *d_kept;
	int64_t n_keep = 0;
	WSOK(wsget(ps, W_UNG, n_dg, &d_ung));
	WSOK(wsget(ps, W_QS, n_dg, &d_qs));
	WSOK(wsget(ps, W_QE, n_dg, &d_qe));
	WSOK(wsget(ps, W_SEL, n_dg, &d_sel));
	UngArgs ua;
	ua.pssm = d_pssm, ua.prow = d_prow, ua.plen = d_plen, ua.aa = ps->d_aa, ua.orf_off = ps->d_orf_off;
	ua.orf_len = ps->d_orf_len, ua.dg = d_dg, ua.n = n_dg, ua.D = D, ua.OD = O + D, ua.dbias = maxL;
	ua.ung = d_ung, ua.qs = d_qs, ua.qe = d_qe;
	hipLaunchKernelGGL(ung_kernel, dim3(blocks(n_dg)), dim3(256), 0, s, ua);
	hipLaunchKernelGGL(best_kernel, dim3(blocks(n_dg)), dim3(256), 0, s, d_dg, n_dg, D, d_ung, d_qs, d_qe,
			in.min_ungapped, d_flag);
	WSOK(hipcub::DeviceSelect::Flagged(d_tmp, tmp_bytes, hipcub::CountingInputIterator<int32_t>(0), d_flag, d_sel,
			d_nsel, (int)n_dg, s));
	WSOK(read_i64(ps, d_nsel, &n_keep));
	nsh_count_launch(3);
	st[4] = n_keep;
	if (n_keep == 0) return 1;
	WSOK(wsget(ps, W_KEPT, n_keep, &d_kept));
	hipLaunchKernelGGL(gather_kernel, dim3(blocks(n_keep)), dim3(256), 0, s, d_sel, n_keep, d_dg, d_ung, d_qs, d_qe,
			d_kept);
	Kept *kept;
	WSOK(pinget(ps, P_OUT, (size_t)n_keep * sizeof(Kept), (void**)&kept));
	WSOK(hipMemcpyAsync(kept, d_kept, (size_t)n_keep * sizeof(Kept), hipMemcpyDeviceToHost, s));
	WSOK(hipStreamSynchronize(s));
	nsh_count_launch(1);
	// per profile (keys are profile-major, ORFs increasing): max_seqs by (score desc, ORF asc)
	struct C1 { int32_t prof, ung, qs, qe, diag; int64_t orf; };
	std::vector<C1> cand;
	const uint64_t omask = (1ull << O) - 1, dmask = (1ull << D) - 1;
	for (int64_t i = 0; i < n_keep; ) {
		const int32_t pf = (int32_t)(kept[i].key >> (O + D));
		const int64_t first = (int64_t)cand.size();
		for (; i < n_keep && (int32_t)(kept[i].key >> (O + D)) == pf; ++i) {
			const uint64_t key = kept[i].key;
			cand.push_back({ pf, kept[i].ung, kept[i].qs, kept[i].qe, (int32_t)((int64_t)(key & dmask) - maxL),
				(int64_t)((key >> D) & omask) });
		}
		if ((int64_t)cand.size() - first > in.max_seqs) {   // prescreen.c:466-469
			std::sort(cand.begin() + first, cand.end(), [](const C1 &x, const C1 &y) {
				if (x.ung != y.ung) return x.ung > y.ung;
				return x.orf < y.orf; });
			cand.resize(first + in.max_seqs);
		}
	}
	const int32_t nc = (int32_t)cand.size();
	st[5] = nc;
	if (nc == 0) return 1;
	// banded gapped scores, grouped by the kernel's column count; one pinned block and one copy:
	// ORF residue offsets, PSSM rows (int64) | ORF lengths, diagonals, profile lengths, items (int32)
	std::vector<std::vector<int32_t>> by_c(17);
	const size_t o32 = al64((size_t)nc * 16), gbytes = o32 + (size_t)nc * 16;
	uint8_t *hg;
	WSOK(pinget(ps, P_IN, gbytes, (void**)&hg));
	int64_t *g64 = (int64_t*)hg;
	int32_t *g32 = (int32_t*)(hg + o32);
	for (int32_t i = 0; i < nc; ++i) {
		const C1 &c = cand[i];
		const int w = in.band > 0? 2 * in.band + 1 : ps->aa_len[c.orf];
		if (in.band <= 0 && w > 1024) { nsh_set_error("prescreen: full-matrix ORF of %d aa exceeds 1024", w); return 0; }
		by_c[std::max(1, (w + 63) / 64)].push_back(i);
		g64[i] = ps->aa_off[c.orf], g64[nc + i] = prow[c.prof];
		g32[i] = ps->aa_len[c.orf], g32[nc + i] = c.diag, g32[2 * nc + i] = plen[c.prof];
	}
	{
		int32_t *it = g32 + 3 * (size_t)nc;
		for (int c = 1; c <= 16; ++c) it = std::copy(by_c[c].begin(), by_c[c].end(), it);
	}
	uint8_t *d_g;
	WSOK(wsget(ps, W_GAP, (int64_t)(gbytes + al64((size_t)nc * 4)), &d_g));
	const int64_t *d_g64 = (const int64_t*)d_g;
	const int32_t *d_g32 = (const int32_t*)(d_g + o32);
	int32_t *d_sw = (int32_t*)(d_g + al64(gbytes));
	WSOK(hipMemcpyAsync(d_g, hg, gbytes, hipMemcpyHostToDevice, s));
	int32_t first = 0;
	for (int c = 1; c <= 16; ++c) {
		if (by_c[c].empty()) continue;
		nsh_psg::launch(c, in.band <= 0, d_pssm, 0, d_g64 + nc, d_g32 + 2 * (size_t)nc, ps->d_aa, d_g64, d_g32,
				d_g32 + nc, in.band, in.go, in.ge, d_g32 + 3 * (size_t)nc + first, (int32_t)by_c[c].size(), d_sw, s);
		first += (int32_t)by_c[c].size();
		nsh_count_launch(1);
	}
	WSOK(hipGetLastError());
	int32_t *sw;
	WSOK(pinget(ps, P_OUT, (size_t)nc * 4, (void**)&sw));
	WSOK(hipMemcpyAsync(sw, d_sw, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
	WSOK(hipStreamSynchronize(s));
	std::unordered_map<int64_t, double> ev_of;   // E-value by (profile length, score)
	// E-values and the output records (prescreen.c:472-481)
	for (int32_t i = 0; i < nc; ++i) {
		const C1 &c = cand[i];
		const int64_t ek = (int64_t)plen[c.prof] << 32 | (uint32_t)sw[i];
		auto it = ev_of.find(ek);
		if (it == ev_of.end()) it = ev_of.emplace(ek, evalue((double)sw[i], (double)plen[c.prof], (double)ps->n_aa)).first;
		const double ev = it->second;
		if (in.max_evalue > 0.0 && ev > in.max_evalue) continue;
		nsh_ps_cand_t r;
		const int64_t vs = ps->vs[c.orf];
		r.vid = ps->vid[c.orf], r.vs = vs, r.ve = vs + 3 * (int64_t)ps->aa_len[c.orf], r.ev = ev;
		r.aqs = c.qs, r.aqe = c.qe, r.asc = c.ung, r.avs = vs + 3 * (int64_t)(c.qs + c.diag);
		res[c.prof].push_back(r);
		++st[6];
	}
	return 1;
}

} // namespace

extern "C" int64_t nsh_prescreen_search_batch(nsh_prescreen_t *ps, int32_t n_prof, const int8_t *pssm, const int32_t *L,
		int32_t thr, int32_t kmer_max, int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge,
		double max_evalue, nsh_ps_cand_t *out, int64_t max_out, int32_t *n_res)
{
	if (!ps || n_prof < 0 || (n_prof > 0 && (!L || !n_res)) || kmer_max < 1 || max_seqs < 0 || ge < 0 || go < ge ||
			go > 65535 || band > 511 || max_out < 0 || (max_out > 0 && !out)) {
		nsh_set_error("prescreen: bad search arguments");
		return -1;
	}
	std::vector<int64_t> prow(n_prof + 1, 0);
	for (int32_t p = 0; p < n_prof; ++p) {
		if (L[p] < 0) { nsh_set_error("prescreen: profile %d has length %d", p, L[p]); return -1; }
		prow[p + 1] = prow[p] + L[p];
	}
	if (prow[n_prof] > 0 && !pssm) { nsh_set_error("prescreen: bad search arguments"); return -1; }
	for (int32_t p = 0; p < n_prof; ++p) n_res[p] = 0;
	std::lock_guard<std::mutex> lk(ps->mu);
	for (int64_t &v : ps->stats) v = 0;
	if (ps->n_orf == 0 || n_prof == 0) return 0;
	SearchIn in = { pssm, L, prow.data(), thr, kmer_max, min_ungapped, max_seqs, band, go, ge, max_evalue };
	int64_t total = 0;
	int32_t p0 = 0, chunk = std::min<int32_t>(n_prof, 1024);
	while (p0 < n_prof) {
		const int32_t p1 = std::min(n_prof, p0 + chunk);
		std::vector<std::vector<nsh_ps_cand_t>> res(p1 - p0);
		const int r = search_chunk(ps, in, p0, p1, res);
		if (r == 0) return -1;
		if (r == 2) {
			if (p1 - p0 == 1) { nsh_set_error("prescreen: profile %d generates more than %lld hits or prefixes", p0,
					(long long)kMaxItems); return -1; }
			chunk = std::max(1, (p1 - p0) / 2);
			continue;
		}
		for (int32_t p = p0; p < p1; ++p) {
			for (const nsh_ps_cand_t &c : res[p - p0]) {
				if (total < max_out) out[total] = c;
				++total;
			}
			n_res[p] = (int32_t)res[p - p0].size();
		}
		p0 = p1;
	}
	return total;
}

extern "C" int nsh_prescreen_stats(nsh_prescreen_t *ps, int64_t *st)
{
	if (!ps || !st) { nsh_set_error("prescreen: bad arguments"); return 0; }
	std::lock_guard<std::mutex> lk(ps->mu);
	for (int i = 0; i < 7; ++i) st[i] = ps->stats[i];
	return 1;
}

extern "C" int32_t nsh_prescreen_search(nsh_prescreen_t *ps, const int8_t *pssm, int32_t L, int32_t thr, int32_t kmer_max,
		int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge, double max_evalue,
		nsh_ps_cand_t *out, int32_t max_out)
{
	int32_t n = 0;
	if (L < 0) { nsh_set_error("prescreen: bad search arguments"); return -1; }
	const int64_t r = nsh_prescreen_search_batch(ps, 1, pssm, &L, thr, kmer_max, min_ungapped, max_seqs, band, go, ge,
			max_evalue, out, max_out, &n);
	return r < 0? -1 : (int32_t)r;
}
