// Internal declarations shared by the host sources of libnasw_hip.
#ifndef NSH_INTERNAL_H
#define NSH_INTERNAL_H

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstddef>
#include <thread>
#include <vector>
#include "../../include/nasw-hip.h"

#define NSH_NEG16 (-32768)

// ---- device/runtime state (nasw_api.cpp) ----
int nsh_ensure_device();               // 1 when a GPU is usable (initialises it once)
void nsh_set_error(const char *fmt, ...);
void nsh_count_launch(uint64_t n);
void nsh_add_times(double h2d, double kernel, double d2h);
int32_t nsh_host_stripe_width(int32_t flag, int32_t al); // the reference CPU's stripe width on this host

// f(i) for i in [0, n) on up to n_threads host threads (the calling thread included)
template <class F> void nsh_parallel_for(int64_t n, int32_t n_threads, F f)
{
	if (n <= 0) return;
	const int32_t nt = (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_threads, n));
	if (nt == 1) {
		for (int64_t i = 0; i < n; ++i) f(i);
		return;
	}
	std::atomic<int64_t> next(0);
	auto work = [&] { for (int64_t i; (i = next.fetch_add(1)) < n;) f(i); };
	std::vector<std::thread> th;
	for (int32_t t = 1; t < nt; ++t) th.emplace_back(work);
	work();
	for (auto &x : th) x.join();
}

// with NSH_PROFILE set in the environment, each lap() prints the host time since the last one to stderr
struct nsh_step_clock {
	const char *who;
	bool on;
	double t;
	explicit nsh_step_clock(const char *w);
	void lap(const char *what, int64_t n = -1);
};

// length penalty of the extension DP: (int)(ie_coef * log2approx(d) + .5f) for d >= 2 (nasw-sse.c:341-349, 758-761)
int32_t nsh_ie_penalty(float ie_coef, int32_t d);

// ---- extension DP engine (ext_dp.hip) ----
// one request as the engine sees it; raw strings are borrowed until the plan is created
struct nsh_ext_req {
	const char *ns; int32_t nl;
	const char *as; int32_t al;
	const uint8_t *ss;
	int32_t left, io;
	int32_t npad;     // al rounded up to the reference CPU's stripe width
	int32_t mode = 0; // 1: score-only global DP without extension flags (nasw-sse.c:459, 559)
	const int8_t *pssm = nullptr; // al x asize PSSM rows in query order (ns_pssm_sc, nasw-sse.c:217-221); null: the matrix
	const uint8_t *d_g = nullptr;  // with g_off >= 0: the bases are d_g[g_off ..) in HBM (ns is not read; no ss)
	int64_t g_off = -1;
};
nsh_ext_plan_t *nsh_ext_plan_build(int32_t n, const nsh_ext_req *req, const ns_opt_t *opt);
int nsh_ext_plan_query(nsh_ext_plan_t *p);   // 1 when the last run has completed
// 1 when nsh_ext_plan_build accepts the request; otherwise 0 and the reason in ns_gpu_error
int nsh_ext_req_fits(const nsh_ext_req &q, int32_t k);

// host-side input preparation (prep_host.cpp), used by the ns_ext_prep() building block
void nsh_prep_host(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int left, uint8_t *nas, uint8_t *aas, int8_t *donor, int8_t *acc);

// ---- global DP engine (global_dp.hip) ----
// one queued problem of ns_global_gs16b (NS_F_CIGAR) / ns_global_gs32b (with or without CIGAR)
// semantics; the inputs are copied, so they need not outlive the call. nullptr (and the reason in
// ns_gpu_error) when the problem is outside the engine's limits.
struct nsh_glob_req;
nsh_glob_req *nsh_glob_submit(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t is32);
// waits for q, appends its CIGAR to r (ns_push_cigar from km) and sets nt_len/aa_len/score; frees q.
// 1 on success
int nsh_glob_wait(nsh_glob_req *q, void *km, ns_rst_t *r);
// submit + wait; 1 on success
int nsh_global_run(void *km, const char *ns, int32_t nl, const char *as, int32_t al,
		const ns_opt_t *opt, const uint8_t *ss, int32_t is32, ns_rst_t *r);

// nsh_global_batch where problem k with g_off[k] >= 0 reads its bases from d_g + g_off[k] in HBM
// (ns may then be null; no ss); run on the caller's thread with its own streams and pools
int nsh_global_batch_g(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *ss, const ns_opt_t *opt, int32_t mode,
		const uint8_t *d_g, const int64_t *g_off, int32_t *score, int32_t *n_cigar, uint32_t **cigar);

// oriented bases (codes 0..4) of [st, en) of strand vid = contig << 1 | reverse, the contig's bases
// at nt4 + ctg_off[c] (mp_ntseq_get_by_v, ntseq.c:92-118); en is clipped to the contig. Returns the
// number written, -1 when st is out of range
int64_t nsh_strand_bases(const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len, uint32_t vid, int64_t st,
		int64_t en, uint8_t *out);

// ---- banded gapped stage of the prescreen search (ps_gapped.hip), device-resident items ----
#include <hip/hip_runtime.h>
#include <vector>
namespace nsh_psg {
// d_prow/d_plen (both null, or both set): per item, the first row of its PSSM in d_pssm and its length
void launch(int C, bool full, const int8_t *d_pssm, int32_t L, const int64_t *d_prow, const int32_t *d_plen,
		const uint8_t *d_aa, const int64_t *d_off, const int32_t *d_len, const int32_t *d_diag, int32_t band, int32_t go,
		int32_t ge, const int32_t *d_items, int32_t n_items, int32_t *d_out, hipStream_t s);
}

// ---- ORF runs of a sequence on the device (orf_scan.hip) ----
namespace nsh_orfs {
struct Work {                      // device buffers, grown on demand; release() frees them
	void *p[6] = {};
	size_t n[6] = {};
	int64_t *st = nullptr;         // after find(): first base of each run
	int32_t *len = nullptr;        // after find(): codons of each run
	void release();
};
// the runs of at least min_orf good codons (ps_orf_worker's walk, prescreen.c:101-132) of an
// oriented sequence already on the device, unsorted; waits for stream s. 1 on success.
int find(hipStream_t s, const uint8_t *d_nt, int64_t len, const uint8_t *d_codon, int32_t min_orf, Work &w,
		int64_t *n_run);
}

// ---- --gpu-prefilter candidate selection (gapless_select.hip), device-resident inputs ----
// queues the selection of a mini-batch on stream s (the calibration waits on s); keep and the K
// per query are written by a kernel still in flight when it returns. 1 on success.
namespace nsh_gsel {
int run(hipStream_t s, int32_t n_query, const int32_t *d_off, const int32_t *d_qlen, const int32_t *d_sc,
		const int32_t *d_chn, const uint32_t *d_vid, const int64_t *d_vs, int32_t n_cand, int32_t max_per_query,
		int32_t top_k, int32_t min_per_vid, int32_t adaptive, uint8_t *d_keep, int32_t *d_k, nsh_gapless_calib_t *cal);
}

// copies between the caller's (pageable) host memory and the device; large ones are staged
// through pinned chunks with the DMA overlapping the host side (host_copy.cpp). Both return
// 1 on success; to_host waits for the data, to_device waits when it staged (otherwise the copy
// is queued on s like hipMemcpyAsync).
int nsh_copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s);
int nsh_copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t s);

namespace nsh_chain {
struct Params {                    // mp_chain's arguments (chain.c:160)
	int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc;
	float coef;
	int32_t is_spliced, kmer, bbit;
};
// mp_chain of every query's sorted anchors d_a[d_off[q], d_off[q] + n_q) on stream s, where the
// count is h_cnt[q] on the host and d_cnt[q] on the device (both null: the offsets' differences,
// d_off holding n_q + 1 entries). Per query, at the same offsets, the chains (score << 32 | count)
// go to d_u and their anchors to d_ao; the chain counts to d_nu. Waits for s; 1 on success.
int run(hipStream_t s, int32_t n_q, const int64_t *d_off, const int64_t *d_cnt, const int64_t *h_cnt,
		const uint64_t *d_a, int64_t n_a, const Params &p, int32_t *d_nu, uint64_t *d_u, uint64_t *d_ao);
}

// ---- a genome in both orientations, on the host and in HBM (the map path's windows) ----
struct nsh_genome_view {
	int32_t n_ctg = 0;
	const int64_t *off = nullptr, *len = nullptr;   // contig c: its bases at off[c] (packed), len[c] of them
	int64_t total = 0;                              // bases of all contigs
	const uint8_t *h_g = nullptr;   // host: the contigs (codes 0..4), then their reverse complements at total + off[c]
	const uint8_t *d_g = nullptr;   // the same in HBM
	// where base pos of strand vid (contig << 1 | reverse, oriented coordinates) is
	int64_t at(uint32_t vid, int64_t pos) const { return (vid & 1? total : 0) + off[vid >> 1] + pos; }
};
// nsh_align_stage over a genome view (align_stage.cpp)
int nsh_align_stage_view(const nsh_genome_view &g, const char *const *aa, const int32_t *qlen, int32_t n_reg,
		nsh_align_reg_t *reg, const uint64_t *anchors, const nsh_align_opt_t *opt, uint32_t **cigar);

// ---- the refinement step of the map path on the device (refine.hip) ----
namespace nsh_refine {
struct Job { int64_t g_off; int32_t len, q; };   // a window: its first base in the oriented genome, its query
// the oriented genome in HBM: the n_ctg contigs (codes 0..4, contig c at ctg_off[c], total bases)
// followed by their reverse complements at total + ctg_off[c]; *d_g is hipFree'd by the caller
int oriented_genome(const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len, int32_t n_ctg, int64_t total,
		uint8_t **d_g);
// mp_refine_reg's sketch, bucket pairs, mp_chain and chain choice (map.c:153-247) for every job. The
// n_q queries' residues are qs[q_off[q] .. q_off[q] + qlen[q]) (n_res bytes in all); their k-mers
// (mp_sketch_prot with kmer and mod_bit 0, sorted: hash << 32 | position) are made on the device.
// Job j's chosen chain is best_u[j] (score << 32 | count; 0 when the job has no chain) and its anchors
// (window position << 32 | query position) are anchors[best_off[j] ..). *n_pairs: the anchor pairs
// chained. 1 on success.
int run(const uint8_t *d_g, const std::vector<Job> &jobs, const char *qs, const int64_t *q_off, const int32_t *qlen,
		int32_t n_q, int64_t n_res, int32_t kmer, int32_t min_aa_len, int32_t max_ava, const nsh_chain::Params &cp,
		std::vector<uint64_t> &best_u, std::vector<int64_t> &best_off, std::vector<uint64_t> &anchors, int64_t *n_pairs);
}

#endif
