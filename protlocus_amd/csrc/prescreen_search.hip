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
// Search (ps_query_worker):
//   * each profile row's 20 residues are ranked by score, stably, as the reference's insertion
//     sort does;
//   * one thread per query window runs the reference's branch-and-bound enumeration of k-mers
//     scoring at least thr, in the same order and with the same cap. The hits are counted in a
//     first pass, then written at scanned offsets in a second;
//   * a device radix sort orders the hits (orf << 32 | diagonal + 2^30);
//   * runs of at least two equal hits are the double-hit diagonals, compacted in order;
//   * one lane per diagonal computes ps_ungapped;
//   * per ORF, the first strictly best diagonal is kept when it reaches min_ungapped.
// The short candidate list is then truncated to max_seqs by (score desc, ORF asc) when it is
// longer. The banded gapped scores come from ps_gapped.hip's kernel on the resident residues.
// The E-value is a few double operations per candidate with the C library's exp/erfc; it is
// evaluated on the host so that its bits equal the reference's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
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

// per profile row: the 20 residues by decreasing score, ties in residue order (stable insertion)
__global__ void rank_kernel(const int8_t *pssm, int32_t L, int8_t *sc, uint8_t *ai)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= L) return;
	int8_t v[PS_AA];
	uint8_t r[PS_AA];
	for (int a = 0; a < PS_AA; ++a) {
		const int8_t x = pssm[(int64_t)i * 22 + a];
		int b = a;
		for (; b > 0 && v[b - 1] < x; --b) v[b] = v[b - 1], r[b] = r[b - 1];
		v[b] = x, r[b] = (uint8_t)a;
	}
	for (int a = 0; a < PS_AA; ++a) sc[(int64_t)i * PS_AA + a] = v[a], ai[(int64_t)i * PS_AA + a] = r[a];
}

struct GenArgs {
	const int8_t *sc; const uint8_t *ai;
	int32_t n_win, thr, cap;
	Pat p;
	const int64_t *off; const uint64_t *ent;
	int64_t *wcount; const int64_t *woff; uint64_t *hit;
};

// one thread per query window q0: the k-mers scoring >= thr by depth-first enumeration over the
// ranked residues (branches cut when even the best completion misses thr), at most cap of them;
// PASS 0 counts their index entries, PASS 1 writes the hits
template <int PASS>
__global__ void __launch_bounds__(64) gen_kernel(GenArgs a)
{
	const int q0 = blockIdx.x * blockDim.x + threadIdx.x;
	if (q0 >= a.n_win) return;
	const int k = a.p.kmer;
	int rest[8], idx[8], sum[8], base[8];
	uint32_t code[8], mul[8];
	for (int t = 0; t < k; ++t) base[t] = (q0 + a.p.pos[t]) * PS_AA;
	rest[k - 1] = 0;
	for (int t = k - 1; t > 0; --t) rest[t - 1] = rest[t] + a.sc[base[t]];
	mul[0] = 1;
	for (int t = 0; t + 1 < k; ++t) mul[t + 1] = mul[t] * PS_AA;
	int64_t n = 0, out = PASS? a.woff[q0] : 0;
	int n_gen = 0, t = 0;
	idx[0] = 0, sum[0] = 0, code[0] = 0;
	while (t >= 0) {
		if (idx[t] >= PS_AA || sum[t] + a.sc[base[t] + idx[t]] + rest[t] < a.thr) {
			--t;
			if (t >= 0) ++idx[t];
			continue;
		}
		if (t + 1 == k) {   // a complete k-mer
			const uint32_t c = code[t] + a.ai[base[t] + idx[t]] * mul[t];
			const int64_t b0 = a.off[c], b1 = a.off[c + 1];
			if (PASS == 0) n += b1 - b0;
			else
				for (int64_t e = b0; e < b1; ++e) {
					const uint64_t x = a.ent[e];
					const int64_t d = (int64_t)(uint32_t)x - q0 + DIAG_OFF;
					a.hit[out++] = (x & 0xffffffff00000000ull) | (uint32_t)d;
				}
			if (++n_gen >= a.cap) break;
			++idx[t];
		} else {
			sum[t + 1] = sum[t] + a.sc[base[t] + idx[t]];
			code[t + 1] = code[t] + a.ai[base[t] + idx[t]] * mul[t];
			++t, idx[t] = 0;
		}
	}
	if (PASS == 0) a.wcount[q0] = n;
}

// first element of each run of >= 2 equal hits: a double-hit diagonal
__global__ void dbl_flag_kernel(const uint64_t *h, int64_t n, uint8_t *f)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	f[k] = k + 1 < n && h[k] == h[k + 1] && (k == 0 || h[k - 1] != h[k]);
}

struct UngArgs {
	const int8_t *pssm; int32_t L;
	const uint8_t *aa; const int64_t *orf_off; const int32_t *orf_len;
	const uint64_t *dg; int64_t n;
	int32_t *ung, *qs, *qe;
};

// ps_ungapped (prescreen.c:334-344) of each double-hit diagonal
__global__ void ung_kernel(UngArgs a)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= a.n) return;
	const int64_t o = (int64_t)(a.dg[k] >> 32);
	const int32_t diag = (int32_t)((uint32_t)a.dg[k] - DIAG_OFF), n = a.orf_len[o];
	const uint8_t *orf = a.aa + a.orf_off[o];
	const int32_t i0 = diag < 0? -diag : 0, i1 = n - diag < a.L? n - diag : a.L;
	int32_t sc = 0, best = 0, st = i0, qs = i0, qe = i0;
	for (int32_t i = i0; i < i1; ++i) {
		sc += a.pssm[(int64_t)i * 22 + orf[i + diag]];
		if (sc < 0) sc = 0, st = i + 1;
		else if (sc > best) best = sc, qs = st, qe = i + 1;
	}
	a.ung[k] = best, a.qs[k] = qs, a.qe[k] = qe;
}

// per ORF (its diagonals are consecutive, in increasing order): the first strictly best diagonal;
// keep[k] marks the ORF's first diagonal when that best reaches min_ungapped
__global__ void best_kernel(const uint64_t *dg, int64_t n, int32_t *ung, int32_t *qs, int32_t *qe, int32_t min_ungapped,
		uint8_t *keep)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const uint64_t o = dg[k] >> 32;
	if (k > 0 && (dg[k - 1] >> 32) == o) { keep[k] = 0; return; }
	int32_t best = -1;
	int64_t bj = k;
	for (int64_t j = k; j < n && (dg[j] >> 32) == o; ++j)
		if (ung[j] > best) best = ung[j], bj = j;
	keep[k] = best >= min_ungapped;
	if (bj != k) {   // move the winner into the ORF's first slot (only this thread touches the run)
		const uint64_t d = dg[bj];
		((uint64_t*)dg)[k] = d;
		ung[k] = ung[bj], qs[k] = qs[bj], qe[k] = qe[bj];
	}
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
	hipStream_t stream = nullptr;
	std::mutex mu;
};

#define PSOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("prescreen: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" void nsh_prescreen_destroy(nsh_prescreen_t *ps)
{
	if (!ps) return;
	if (ps->stream) (void)hipStreamSynchronize(ps->stream);
	void *bufs[] = { ps->d_aa, ps->d_orf_off, ps->d_off, ps->d_orf_len, ps->d_ent };
	for (void *b : bufs) if (b) (void)hipFree(b);
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

extern "C" int32_t nsh_prescreen_search(nsh_prescreen_t *ps, const int8_t *pssm, int32_t L, int32_t thr, int32_t kmer_max,
		int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge, double max_evalue,
		nsh_ps_cand_t *out, int32_t max_out)
{
	if (!ps || (L > 0 && !pssm) || L < 0 || kmer_max < 1 || max_seqs < 0 || ge < 0 || go < ge || go > 65535 ||
			band > 511 || max_out < 0 || (max_out > 0 && !out)) {
		nsh_set_error("prescreen: bad search arguments");
		return -1;
	}
	if (L < ps->pat.span || ps->n_orf == 0) return 0;
	std::lock_guard<std::mutex> lk(ps->mu);
	const int32_t n_win = L - ps->pat.span + 1;
	int8_t *d_pssm = nullptr, *d_sc = nullptr;
	uint8_t *d_ai = nullptr, *d_flag = nullptr;
	int64_t *d_wcount = nullptr, *d_woff = nullptr, *d_nsel = nullptr;
	uint64_t *d_hit = nullptr, *d_hit2 = nullptr, *d_dg = nullptr;
	int32_t *d_ung = nullptr, *d_qs = nullptr, *d_qe = nullptr, *d_items = nullptr, *d_diag = nullptr, *d_sw = nullptr;
	void *d_tmp = nullptr;
	size_t tmp_bytes = 0, need = 0;
	int64_t n_hit = 0, n_dg = 0;
	int32_t n_res = -1;
	std::vector<int64_t> wc;
	{
		hipStream_t s = ps->stream;
		PSOK(hipMalloc((void**)&d_pssm, (size_t)L * 22 + 8));
		PSOK(hipMalloc((void**)&d_sc, (size_t)L * PS_AA + 8));
		PSOK(hipMalloc((void**)&d_ai, (size_t)L * PS_AA + 8));
		PSOK(hipMalloc((void**)&d_wcount, (size_t)n_win * 8 + 8));
		PSOK(hipMalloc((void**)&d_woff, (size_t)(n_win + 1) * 8));
		PSOK(hipMemcpyAsync(d_pssm, pssm, (size_t)L * 22, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(rank_kernel, dim3((L + 63) / 64), dim3(64), 0, s, d_pssm, L, d_sc, d_ai);
		GenArgs ga;
		ga.sc = d_sc, ga.ai = d_ai, ga.n_win = n_win, ga.thr = thr, ga.cap = kmer_max, ga.p = ps->pat;
		ga.off = ps->d_off, ga.ent = ps->d_ent, ga.wcount = d_wcount, ga.woff = d_woff, ga.hit = nullptr;
		hipLaunchKernelGGL(gen_kernel<0>, dim3((n_win + 63) / 64), dim3(64), 0, s, ga);
		PSOK(hipGetLastError());
		wc.resize(n_win + 1);
		PSOK(hipMemcpyAsync(wc.data(), d_wcount, (size_t)n_win * 8, hipMemcpyDeviceToHost, s));
		PSOK(hipStreamSynchronize(s));
		{   // window offsets (one int per window)
			int64_t acc = 0;
			for (int32_t q = 0; q < n_win; ++q) { const int64_t c = wc[q]; wc[q] = acc; acc += c; }
			wc[n_win] = acc;
			n_hit = acc;
		}
		if (n_hit == 0) { n_res = 0; goto done; }
		if (n_hit > INT_MAX) { nsh_set_error("prescreen: %lld hits exceed one sort", (long long)n_hit); goto fail; }
		PSOK(hipMemcpyAsync(d_woff, wc.data(), (size_t)(n_win + 1) * 8, hipMemcpyHostToDevice, s));
		PSOK(hipMalloc((void**)&d_hit, (size_t)n_hit * 8));
		PSOK(hipMalloc((void**)&d_hit2, (size_t)n_hit * 8));
		ga.hit = d_hit;
		hipLaunchKernelGGL(gen_kernel<1>, dim3((n_win + 63) / 64), dim3(64), 0, s, ga);
		PSOK(hipGetLastError());
		// sort, then compact the double-hit diagonals (order kept)
		PSOK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, d_hit, d_hit2, (int)n_hit, 0, 64, s));
		PSOK(hipMalloc((void**)&d_flag, (size_t)n_hit + 8));
		PSOK(hipMalloc((void**)&d_dg, (size_t)n_hit * 8));
		PSOK(hipMalloc((void**)&d_nsel, 8));
		PSOK(hipcub::DeviceSelect::Flagged(nullptr, need, d_hit2, d_flag, d_dg, d_nsel, (int)n_hit, s));
		tmp_bytes = std::max(tmp_bytes, need);
		PSOK(hipMalloc(&d_tmp, tmp_bytes));
		PSOK(hipcub::DeviceRadixSort::SortKeys(d_tmp, tmp_bytes, d_hit, d_hit2, (int)n_hit, 0, 64, s));
		hipLaunchKernelGGL(dbl_flag_kernel, dim3((unsigned)((n_hit + 255) / 256)), dim3(256), 0, s, d_hit2, n_hit, d_flag);
		PSOK(hipcub::DeviceSelect::Flagged(d_tmp, tmp_bytes, d_hit2, d_flag, d_dg, d_nsel, (int)n_hit, s));
		PSOK(hipMemcpyAsync(&n_dg, d_nsel, 8, hipMemcpyDeviceToHost, s));
		PSOK(hipStreamSynchronize(s));
		if (n_dg == 0) { n_res = 0; goto done; }
		PSOK(hipMalloc((void**)&d_ung, (size_t)n_dg * 4));
		PSOK(hipMalloc((void**)&d_qs, (size_t)n_dg * 4));
		PSOK(hipMalloc((void**)&d_qe, (size_t)n_dg * 4));
		UngArgs ua;
		ua.pssm = d_pssm, ua.L = L, ua.aa = ps->d_aa, ua.orf_off = ps->d_orf_off, ua.orf_len = ps->d_orf_len;
		ua.dg = d_dg, ua.n = n_dg, ua.ung = d_ung, ua.qs = d_qs, ua.qe = d_qe;
		hipLaunchKernelGGL(ung_kernel, dim3((unsigned)((n_dg + 255) / 256)), dim3(256), 0, s, ua);
		hipLaunchKernelGGL(best_kernel, dim3((unsigned)((n_dg + 255) / 256)), dim3(256), 0, s, d_dg, n_dg, d_ung, d_qs, d_qe,
				min_ungapped, d_flag);
		PSOK(hipGetLastError());
		// the per-ORF candidates (one per kept ORF, in ORF order): a short list, finished on the host
		std::vector<uint8_t> keep(n_dg);
		std::vector<uint64_t> dg(n_dg);
		std::vector<int32_t> ung(n_dg), qs(n_dg), qe(n_dg);
		PSOK(hipMemcpyAsync(keep.data(), d_flag, (size_t)n_dg, hipMemcpyDeviceToHost, s));
		PSOK(hipMemcpyAsync(dg.data(), d_dg, (size_t)n_dg * 8, hipMemcpyDeviceToHost, s));
		PSOK(hipMemcpyAsync(ung.data(), d_ung, (size_t)n_dg * 4, hipMemcpyDeviceToHost, s));
		PSOK(hipMemcpyAsync(qs.data(), d_qs, (size_t)n_dg * 4, hipMemcpyDeviceToHost, s));
		PSOK(hipMemcpyAsync(qe.data(), d_qe, (size_t)n_dg * 4, hipMemcpyDeviceToHost, s));
		PSOK(hipStreamSynchronize(s));
		std::vector<int64_t> cand;
		for (int64_t k = 0; k < n_dg; ++k) if (keep[k]) cand.push_back(k);
		if ((int64_t)cand.size() > max_seqs) {   // prescreen.c:466-469: best ungapped first, ties by ORF
			std::sort(cand.begin(), cand.end(), [&](int64_t x, int64_t y) {
				if (ung[x] != ung[y]) return ung[x] > ung[y];
				return (dg[x] >> 32) < (dg[y] >> 32); });
			cand.resize(max_seqs);
		}
		const int32_t nc = (int32_t)cand.size();
		if (nc == 0) { n_res = 0; goto done; }
		// banded gapped scores of the candidates, grouped by the kernel's column count
		std::vector<int32_t> diag(nc), items, sw(nc);
		std::vector<std::vector<int32_t>> by_c(17);
		for (int32_t i = 0; i < nc; ++i) {
			const int64_t o = (int64_t)(dg[cand[i]] >> 32);
			diag[i] = (int32_t)((uint32_t)dg[cand[i]] - DIAG_OFF);
			const int w = band > 0? 2 * band + 1 : ps->aa_len[o];
			if (band <= 0 && w > 1024) { nsh_set_error("prescreen: full-matrix ORF of %d aa exceeds 1024", w); goto fail; }
			by_c[std::max(1, (w + 63) / 64)].push_back(i);
		}
		for (int c = 1; c <= 16; ++c) items.insert(items.end(), by_c[c].begin(), by_c[c].end());
		{
			std::vector<int64_t> coff(nc);
			std::vector<int32_t> clen(nc);
			for (int32_t i = 0; i < nc; ++i) {
				const int64_t o = (int64_t)(dg[cand[i]] >> 32);
				coff[i] = ps->aa_off[o], clen[i] = ps->aa_len[o];
			}
			int64_t *d_coff = nullptr;
			int32_t *d_clen = nullptr;
			PSOK(hipMalloc((void**)&d_items, (size_t)nc * 4 * 4 + 8));   // items, diag, length, scores
			d_diag = d_items + nc, d_clen = d_items + 2 * (size_t)nc, d_sw = d_items + 3 * (size_t)nc;
			PSOK(hipMalloc((void**)&d_coff, (size_t)nc * 8));
			PSOK(hipMemcpyAsync(d_items, items.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
			PSOK(hipMemcpyAsync(d_diag, diag.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
			PSOK(hipMemcpyAsync(d_clen, clen.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
			PSOK(hipMemcpyAsync(d_coff, coff.data(), (size_t)nc * 8, hipMemcpyHostToDevice, s));
			int32_t first = 0;
			for (int c = 1; c <= 16; ++c) {
				if (by_c[c].empty()) continue;
				nsh_psg::launch(c, band <= 0, d_pssm, L, ps->d_aa, d_coff, d_clen, d_diag, band, go, ge, d_items + first,
						(int32_t)by_c[c].size(), d_sw, s);
				first += (int32_t)by_c[c].size();
			}
			const hipError_t e1 = hipGetLastError();
			const hipError_t e2 = hipMemcpyAsync(sw.data(), d_sw, (size_t)nc * 4, hipMemcpyDeviceToHost, s);
			const hipError_t e3 = hipStreamSynchronize(s);
			(void)hipFree(d_coff);
			PSOK(e1); PSOK(e2); PSOK(e3);
		}
		// E-values and the output records (prescreen.c:472-481)
		n_res = 0;
		for (int32_t i = 0; i < nc; ++i) {
			const int64_t k = cand[i], o = (int64_t)(dg[k] >> 32);
			const double ev = evalue((double)sw[i], (double)L, (double)ps->n_aa);
			if (max_evalue > 0.0 && ev > max_evalue) continue;
			if (n_res < max_out) {
				nsh_ps_cand_t &r = out[n_res];
				r.vid = ps->vid[o], r.vs = ps->vs[o], r.ve = ps->vs[o] + 3 * (int64_t)ps->aa_len[o], r.ev = ev;
				r.aqs = qs[k], r.aqe = qe[k], r.asc = ung[k], r.avs = ps->vs[o] + 3 * (int64_t)(qs[k] + diag[i]);
			}
			++n_res;
		}
		nsh_count_launch(8);
	}
done:
fail:
	{
		void *bufs[] = { d_pssm, d_sc, d_ai, d_flag, d_wcount, d_woff, d_nsel, d_hit, d_hit2, d_dg, d_ung, d_qs, d_qe,
			d_items, d_tmp };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	return n_res;
}
