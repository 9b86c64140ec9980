// The align stage of the map path for many candidate regions at once, on the device engines
// (nsh_align_stage; the reference runs it per query through mp_align_batch, align.c:745-777).
//
// Each candidate goes through the stages of mp_align_step (align.c:648-743) with the reference's
// default options: trusted-anchor filter, left extension (and its io_end retry for a 5'-end exon),
// the gap DPs between trusted anchors, right extension (and its retry), the DP over the right
// extension, then the alignment's statistics and records. Here every stage runs for all candidates
// together: one extension plan per extension stage, one global-DP engine batch per DP stage, so a
// stage costs one round trip to the GPU however many proteins are in flight. The host only does
// what the reference does between the DPs: interval bookkeeping, the ungapped scores of anchor-to-
// anchor steps and CIGAR concatenation.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr uint64_t kTrusted = 1ULL << 31;   // mark of a trusted anchor in the query field

int32_t anchor_q(uint64_t a) { return (int32_t)((uint32_t)a << 1) >> 1; }   // query field without the mark
int64_t anchor_x(uint64_t a) { return (int64_t)(a >> 32); }

// mp_filter_seed (align.c:8-34): anchors in runs of at least min_cnt whose steps are whole codons,
// at most max_aa_dist residues long, are trusted, except those within trim_back residues of the
// run's ends
void trust_runs(int32_t cnt, uint64_t *a, int32_t max_aa_dist, int32_t min_cnt, int32_t kmer2, int32_t trim_back)
{
	for (int32_t i = 0; i < cnt; ++i) {
		int32_t j = i + 1;
		for (; j < cnt; ++j) {
			const int32_t dx = (int32_t)(a[j] >> 32) - (int32_t)(a[j - 1] >> 32), dy = (int32_t)a[j] - (int32_t)a[j - 1];
			if (dx % 3 != 0 || dx > max_aa_dist * 3 || dy > max_aa_dist) break;
		}
		if (j - i < min_cnt) continue;
		int32_t k = j - 2;
		const int32_t last = (int32_t)a[j - 1];
		while (k >= i && last - (int32_t)a[k] < trim_back) --k;
		const int32_t first = (int32_t)a[i] + 1 - kmer2;
		while (i < k && (int32_t)a[i] + 1 - first < trim_back) ++i;
		for (; i <= k; ++i) a[i] |= kTrusted;
		i = j - 1;
	}
}

// the op appended as ns_push_cigar does (nasw.h:219-229): F and G are never merged
void push_op(std::vector<uint32_t> &c, uint32_t op, int32_t len)
{
	if (c.empty() || op != (c.back() & 0xf) || op == NS_CIGAR_F || op == NS_CIGAR_G) c.push_back((uint32_t)len << 4 | op);
	else c.back() += (uint32_t)len << 4;
}

struct Cand {
	nsh_align_reg_t *r = nullptr;
	const char *aa = nullptr;
	int32_t len = 0;
	std::vector<uint64_t> a;
	const uint8_t *nt = nullptr; // oriented bases of [as, ae) in the host genome view
	int64_t g_off = 0;           // where they start in the oriented genome (host and HBM alike)
	const uint8_t *d_g = nullptr; // the oriented genome in HBM
	int64_t as = 0, ae = 0, vs0 = 0, vs1 = 0;
	int32_t i0 = 0, ne0 = 0, ae0 = 0, score = 0;
	int32_t as1 = 0, as1_win = 0, aa_win = 0, nt_len = 0, aa_len = 0;
	int64_t ve1 = 0;             // end of the last trusted anchor: where the right extension starts
	int32_t qe1 = 0;
	int32_t rnt_len = 0, raa_len = 0;   // the right extension's result
	bool alive = false, right = false;
	std::vector<uint32_t> cigar;
};

// one DP of mp_align_seq (align.c:103-145): an anchor-to-anchor step of exactly alen codons and at
// most kmer2 residues is scored without a DP; anything else goes to the global DP with CIGAR
struct Task {
	Task() {}             // every field is written where a task is made (no zero fill of millions)
	Cand *c;
	int64_t nt_off;
	int32_t nlen, aa_off, alen;
	int32_t score;
	int64_t cig_off;      // a DP's CIGAR: n_cig ops at cig_off of its round's CIGAR array
	int32_t n_cig;
	bool dp;
};
// a stage's DP tasks, candidate after candidate, and the CIGARs of its DPs
struct TaskSet {
	std::vector<Task> t;
	std::vector<int64_t> first;   // tasks of candidate k: t[first[k] .. first[k + 1])
	std::vector<uint32_t> cig;
};

}   // namespace

// oriented bases of [st, en) of strand vid (mp_ntseq_get_by_v, ntseq.c:92-118); -1 when st is out of range
int64_t nsh_strand_bases(const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len, uint32_t vid, int64_t st,
		int64_t en, uint8_t *out)
{
	const int64_t L = ctg_len[vid >> 1], base = ctg_off[vid >> 1];
	if (st < 0 || en < 0 || st >= L) return -1;
	if (en > L) en = L;
	int64_t k = 0;
	if (!(vid & 1)) {
		for (int64_t i = st; i < en; ++i) out[k++] = nt4[base + i];
	} else {   // reverse strand: complement of the forward bases, right to left
		for (int64_t i = L - st - 1; i >= L - en; --i) {
			const uint8_t c = nt4[base + i];
			out[k++] = c >= 4? c : (uint8_t)(3 - c);
		}
	}
	return k;
}

extern "C" void nsh_align_opt_default(nsh_align_opt_t *o)
{
	// mp_mapopt_init (options.c:42-138) as main() leaves it without -I / -G
	std::memset(o, 0, sizeof(*o));
	o->go = 11, o->ge = 1, o->io = 29, o->io_end = 19, o->fs = 23, o->xdrop = 100, o->end_bonus = 5;
	o->max_ext = 10000, o->max_intron = 200000, o->kmer2 = 5, o->end_retry_win = 10000;
	o->ie_coef = .5f, o->sp_scale = 1.0f, o->sp_model = NS_S_GENERIC, o->sp_null_bonus = -7;
	std::memcpy(o->mat, ns_mat_blosum62, 484);
	ns_set_stop_sc(22, o->mat, (int8_t)o->fs);
}

// with NSH_PROFILE set in the environment, the host time of each step goes to stderr
struct StepClock {
	bool on = std::getenv("NSH_PROFILE") != nullptr;
	std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
	void lap(const char *what, int64_t n)
	{
		if (!on) return;
		const auto now = std::chrono::steady_clock::now();
		std::fprintf(stderr, "[nsh_align_stage] %-22s %9.1f ms  n=%lld\n", what,
				std::chrono::duration<double, std::milli>(now - t).count(), (long long)n);
		t = now;
	}
};

static int ext_round(std::vector<Cand*> &cs, std::vector<nsh_ext_req> &req, const ns_opt_t &o, std::vector<nsh_ext_result_t> &res)
{
	res.assign(req.size(), nsh_ext_result_t{});
	if (req.empty()) return 1;
	StepClock clk;
	if (clk.on) {   // the batch's shape: rows, rows x padded columns, and its longest problems
		int64_t rows = 0, cells = 0, nl_max = 0, n_long = 0, rows_long = 0;
		for (const nsh_ext_req &q : req) {
			rows += q.nl, cells += (int64_t)q.nl * q.npad, nl_max = std::max<int64_t>(nl_max, q.nl);
			if (q.nl >= 20000) ++n_long, rows_long += q.nl;
		}
		std::fprintf(stderr, "[nsh_align_stage]   batch: %zu problems, %lld rows, %.3g padded cells, longest %lld rows, "
				"%lld of >= 20,000 rows (%lld rows)\n", req.size(), (long long)rows, (double)cells, (long long)nl_max,
				(long long)n_long, (long long)rows_long);
		std::vector<std::pair<int64_t, int32_t>> big;
		for (int32_t k = 0; k < (int32_t)req.size(); ++k) big.push_back({(int64_t)req[k].nl * ((req[k].npad + 127) / 128), k});
		std::sort(big.begin(), big.end(), std::greater<std::pair<int64_t, int32_t>>());
		int64_t hist[5] = {};
		for (const nsh_ext_req &q : req) hist[std::min(4, (q.npad - 1) / 128)] += q.nl;
		std::fprintf(stderr, "[nsh_align_stage]   rows by padded width <=128/256/384/512/more: %lld %lld %lld %lld %lld; costliest (nl, al):",
				(long long)hist[0], (long long)hist[1], (long long)hist[2], (long long)hist[3], (long long)hist[4]);
		for (size_t t = 0; t < big.size() && t < 8; ++t) std::fprintf(stderr, " (%d, %d)", req[big[t].second].nl, req[big[t].second].al);
		std::fprintf(stderr, "\n");
	}
	nsh_ext_plan_t *p = nsh_ext_plan_build((int32_t)req.size(), req.data(), &o);
	clk.lap("  plan build", (int64_t)req.size());
	const int ok = p && nsh_ext_plan_run(p) && nsh_ext_plan_results(p, res.data());
	clk.lap("  plan run", (int64_t)req.size());
	if (p) nsh_ext_plan_destroy(p);
	(void)cs;
	return ok;
}

static nsh_ext_req ext_req(const Cand *c, int64_t nt_off, int32_t nl, int32_t aa_off, int32_t al, int32_t left, int32_t io)
{
	nsh_ext_req q;
	q.ns = (const char*)c->nt + nt_off, q.nl = nl, q.as = c->aa + aa_off, q.al = al, q.ss = nullptr;
	q.d_g = c->d_g, q.g_off = c->g_off + nt_off;   // the extension reads its bases from the genome in HBM
	q.left = left, q.io = io, q.mode = 0, q.pssm = nullptr;
	const int32_t w = nsh_host_stripe_width(left? NS_F_EXT_LEFT : NS_F_EXT_RIGHT, al);
	q.npad = (al + w - 1) / w * w;
	return q;
}

// the DPs of a stage: ungapped steps on the host threads, the others as one global-DP batch whose
// bases stay in HBM
static int dp_round(TaskSet &ts, const ns_opt_t &o, const nsh_align_opt_t *ao, int32_t nth)
{
	StepClock clk;
	std::vector<Task> &tasks = ts.t;
	const int64_t nt = (int64_t)tasks.size();
	nsh_parallel_for((nt + 4095) / 4096, nth, [&](int64_t b) {
		for (int64_t t = b * 4096; t < std::min(nt, (b + 1) * 4096); ++t) {
			Task &k = tasks[t];
			const Cand *c = k.c;
			k.dp = !(k.nlen == k.alen * 3 && k.alen <= ao->kmer2);
			k.n_cig = 0, k.cig_off = 0;
			if (k.dp) continue;
			// mp_score_ungapped (align.c:36-46): codons with an N score as X. Its loop bound is the
			// residue count while its index steps over bases (i < alen, i += 3), so only the first
			// ceil(alen / 3) residues of the step are scored; the same bound is kept here.
			int32_t sc = 0;
			const uint8_t *nseq = c->nt + k.nt_off;
			for (int32_t i = 0, j = 0; i < k.alen; i += 3, ++j) {
				const uint8_t b0 = nseq[i], b1 = nseq[i + 1], b2 = nseq[i + 2];
				const int32_t ta = b0 > 3 || b1 > 3 || b2 > 3? ns_tab_aa20['X'] : ns_tab_codon[b0 << 4 | b1 << 2 | b2];
				sc += ao->mat[ta * 22 + ns_tab_aa20[(uint8_t)c->aa[k.aa_off + j]]];
			}
			k.score = sc;
		}
	});
	std::vector<int64_t> idx;
	for (int64_t t = 0; t < nt; ++t) if (tasks[t].dp) idx.push_back(t);
	clk.lap("  ungapped steps", nt - (int64_t)idx.size());
	ts.cig.clear();
	if (idx.empty()) return 1;
	const int32_t n = (int32_t)idx.size();
	// the bases stay in HBM (offsets into the oriented genome); only the protein pieces are packed
	std::vector<int64_t> ns_off(n, 0), g_off(n), as_off(n);
	std::vector<int32_t> nl(n), al(n), score(n), ncig(n);
	int64_t l_as = 0;
	for (int32_t k = 0; k < n; ++k) {
		const Task &t = tasks[idx[k]];
		g_off[k] = t.c->g_off + t.nt_off, as_off[k] = l_as, l_as += t.alen;
		nl[k] = t.nlen, al[k] = t.alen;
	}
	std::vector<char> as((size_t)l_as + 1);
	nsh_parallel_for((n + 4095) / 4096, nth, [&](int64_t b) {
		for (int32_t k = (int32_t)(b * 4096); k < std::min<int64_t>(n, (b + 1) * 4096); ++k) {
			const Task &t = tasks[idx[k]];
			std::memcpy(as.data() + as_off[k], t.c->aa + t.aa_off, (size_t)t.alen);
		}
	});
	clk.lap("  DP inputs", n);
	uint32_t *cig = nullptr;
	ns_opt_t oc = o;
	oc.flag = NS_F_CIGAR;
	// mode 2: ns_global_gs_auto's int16 / int32 choice, as mp_align_seq calls it
	if (!nsh_global_batch_g(n, nullptr, ns_off.data(), nl.data(), as.data(), as_off.data(), al.data(), nullptr, &oc, 2,
			tasks[idx[0]].c->d_g, g_off.data(), score.data(), ncig.data(), &cig))
		return 0;
	clk.lap("  DP batch", n);
	int64_t off = 0;
	for (int32_t k = 0; k < n; ++k) {
		Task &t = tasks[idx[k]];
		t.score = score[k], t.cig_off = off, t.n_cig = ncig[k];
		off += ncig[k];
	}
	ts.cig.assign(cig, cig + off);
	std::free(cig);
	return 1;
}

// each candidate's task results, in its order: scores added, CIGARs appended (ns_push_cigar's merge)
static void merge_tasks(const TaskSet &ts, const std::vector<Cand*> &cs, int32_t nth)
{
	nsh_parallel_for((int64_t)cs.size(), nth, [&](int64_t k) {
		Cand *c = cs[k];
		for (int64_t t = ts.first[k]; t < ts.first[k + 1]; ++t) {
			const Task &x = ts.t[t];
			c->score += x.score;
			if (!x.dp) push_op(c->cigar, NS_CIGAR_M, x.alen);
			else for (int32_t j = 0; j < x.n_cig; ++j) push_op(c->cigar, ts.cig[x.cig_off + j] & 0xf, (int32_t)(ts.cig[x.cig_off + j] >> 4));
		}
	});
}

extern "C" int nsh_align_stage(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		const char *const *aa, const int32_t *qlen, int32_t n_reg, nsh_align_reg_t *reg, const uint64_t *anchors,
		const nsh_align_opt_t *ao, uint32_t **cigar_out)
{
	if (!nt4 || !ctg_off || !ctg_len || n_ctg <= 0 || !cigar_out) {
		nsh_set_error("nsh_align_stage: bad arguments");
		return 0;
	}
	// the oriented genome of this call: the caller's contigs packed one after another, then their
	// reverse complements, on the host and in HBM
	std::vector<int64_t> off(n_ctg);
	int64_t total = 0;
	for (int32_t c = 0; c < n_ctg; ++c) off[c] = total, total += ctg_len[c];
	std::vector<uint8_t> hg((size_t)(2 * total + 1));
	for (int32_t c = 0; c < n_ctg; ++c) {
		std::memcpy(hg.data() + off[c], nt4 + ctg_off[c], (size_t)ctg_len[c]);
		nsh_strand_bases(nt4, ctg_off, ctg_len, (uint32_t)c << 1 | 1, 0, ctg_len[c], hg.data() + total + off[c]);
	}
	uint8_t *d_g = nullptr;
	if (!nsh_refine::oriented_genome(hg.data(), off.data(), ctg_len, n_ctg, total, &d_g)) return 0;
	nsh_genome_view g;
	g.n_ctg = n_ctg, g.off = off.data(), g.len = ctg_len, g.total = total, g.h_g = hg.data(), g.d_g = d_g;
	const int ok = nsh_align_stage_view(g, aa, qlen, n_reg, reg, anchors, ao, cigar_out);
	(void)hipFree(d_g);
	return ok;
}

int nsh_align_stage_view(const nsh_genome_view &g, const char *const *aa, const int32_t *qlen, int32_t n_reg,
		nsh_align_reg_t *reg, const uint64_t *anchors, const nsh_align_opt_t *ao, uint32_t **cigar_out)
{
	if (!g.h_g || !g.d_g || !aa || !qlen || (n_reg > 0 && (!reg || !anchors)) || !ao || !cigar_out || n_reg < 0) {
		nsh_set_error("nsh_align_stage: bad arguments");
		return 0;
	}
	*cigar_out = nullptr;
	if (!nsh_ensure_device()) return 0;
	const int32_t n_ctg = g.n_ctg;
	const int64_t *ctg_len = g.len;
	// ns_opt_t of the DPs (mp_map2ns_opt, align.c:59-71)
	ns_opt_t o;
	ns_opt_init(&o);
	o.go = ao->go, o.ge = ao->ge, o.io = ao->io, o.fs = ao->fs, o.xdrop = ao->xdrop, o.sc = ao->mat;
	o.ie_coef = ao->ie_coef, o.end_bonus = ao->end_bonus, o.sp_null_bonus = ao->sp_null_bonus;
	ns_opt_set_sp(&o, ao->sp_model);
	for (int t = 0; t < 6; ++t) o.sp[t] = (int32_t)(o.sp[t] * ao->sp_scale + .499f);
	StepClock clk;
	std::vector<Cand> cand(n_reg);
	std::vector<Cand*> live;
	const int32_t nth = 16;
	// MP_AS_START: trusted anchors, the search window, its bases (mp_align_start, align.c:431-480),
	// every candidate on the host threads; the first region outside its contig fails the call
	std::atomic<int32_t> bad_reg(INT32_MAX);
	nsh_parallel_for(((int64_t)n_reg + 255) / 256, nth, [&](int64_t b) {
	for (int32_t k = (int32_t)(b * 256); k < std::min<int64_t>(n_reg, (b + 1) * 256); ++k) {
		nsh_align_reg_t *r = &reg[k];
		Cand &c = cand[k];
		c.r = r, r->status = 0, r->dp_score = 0, r->n_cigar = 0, r->cigar_off = 0;
		if ((r->vid >> 1) >= (uint32_t)n_ctg || r->query < 0 || r->cnt <= 0) continue;
		c.aa = aa[r->query], c.len = qlen[r->query];
		c.a.assign(anchors + r->a_off, anchors + r->a_off + r->cnt);
		trust_runs(r->cnt, c.a.data(), 6, 3, ao->kmer2, ao->kmer2 + 1);
		int32_t i = 0;
		while (i < r->cnt && !(c.a[i] & kTrusted)) ++i;
		if (i == r->cnt) { r->cnt = 0; continue; }   // no trusted anchor: the candidate is dropped
		c.i0 = i;
		int32_t extl = ao->max_ext, extr = ao->max_ext;
		if (r->qs >= 10) extl = ao->max_intron / 2;
		if (c.len - r->qe >= 10) extr = ao->max_intron / 2;
		const int32_t extl0 = (int32_t)(r->ext >> 32), extr0 = (int32_t)r->ext;
		if (extl0 > 0) extl = std::min(extl, extl0);
		if (extr0 > 0) extr = std::min(extr, extr0);
		const int64_t clen = ctg_len[r->vid >> 1];
		c.as = r->vs > extl? r->vs - extl : 0;
		c.ae = r->ve + extr < clen? r->ve + extr : clen;
		if (c.as >= clen || c.ae <= c.as) {
			for (int32_t x = bad_reg.load(); k < x && !bad_reg.compare_exchange_weak(x, k);) {}
			continue;
		}
		c.g_off = g.at(r->vid, c.as), c.nt = g.h_g + c.g_off, c.d_g = g.d_g;
		c.vs0 = r->vs;
		c.vs1 = c.vs0 + anchor_x(c.a[i]) + 1;
		c.as1 = anchor_q(c.a[i]) + 1;
		c.as1_win = c.as1;
		c.alive = true;
	}
	});
	if (bad_reg.load() != INT32_MAX) {
		nsh_set_error("nsh_align_stage: region %d outside its contig", bad_reg.load());
		return 0;
	}
	for (Cand &c : cand) if (c.alive) live.push_back(&c);
	clk.lap("start", (int64_t)live.size());
	// MP_AS_LEFT and MP_AS_RIGHT in one extension plan: the right extension starts at the end of the
	// last trusted anchor, which neither the left extension nor the gap DPs move (the gap steps end
	// on the trusted anchors), so both directions run together and the batch has one tail instead
	// of two; their io_end retries (MP_AS_LEFT_RETRY, MP_AS_RIGHT_RETRY) form a second plan
	std::vector<nsh_ext_req> req;
	std::vector<nsh_ext_result_t> res;
	std::vector<Cand*> rlive, who, rwho;
	for (Cand *c : live) {
		int32_t il = c->r->cnt - 1;
		while (!(c->a[il] & kTrusted)) --il;
		c->ve1 = c->vs0 + anchor_x(c->a[il]) + 1, c->qe1 = anchor_q(c->a[il]) + 1;
		if (c->qe1 < c->len && c->ve1 < c->ae) c->aa_win = c->len - c->qe1, c->right = true, rlive.push_back(c);
	}
	for (Cand *c : live) req.push_back(ext_req(c, 0, (int32_t)(c->vs1 - c->as), 0, c->as1_win, 1, ao->io));
	for (Cand *c : rlive) req.push_back(ext_req(c, c->ve1 - c->as, (int32_t)(c->ae - c->ve1), c->qe1, c->aa_win, 0, ao->io));
	if (!ext_round(live, req, o, res)) return 0;
	req.clear();
	const size_t n_live = live.size();
	for (size_t k = 0; k < n_live; ++k) {
		Cand *c = live[k];
		c->nt_len = res[k].nt_len, c->aa_len = res[k].aa_len;
		// a 5'-end exon missed by the extension gets another try at io_end
		if (res[k].aa_len != c->as1_win && res[k].nt_len < ao->max_ext && ao->io > ao->io_end) {
			const int64_t as_alt = c->vs1 - c->as > ao->end_retry_win? c->vs1 - ao->end_retry_win : c->as;
			req.push_back(ext_req(c, as_alt - c->as, (int32_t)(c->vs1 - as_alt), 0, c->as1_win, 1, ao->io_end));
			who.push_back(c);
		}
	}
	for (size_t k = 0; k < rlive.size(); ++k) {
		Cand *c = rlive[k];
		c->rnt_len = res[n_live + k].nt_len, c->raa_len = res[n_live + k].aa_len;
		// the same for a 3'-end exon
		if (c->raa_len < c->aa_win && c->rnt_len < ao->max_ext && ao->io > ao->io_end) {
			const int32_t l_ext = (int32_t)std::min<int64_t>(c->ae - c->ve1, ao->end_retry_win);
			req.push_back(ext_req(c, c->ve1 - c->as, l_ext, c->qe1, c->aa_win, 0, ao->io_end));
			rwho.push_back(c);
		}
	}
	clk.lap("extensions", (int64_t)(n_live + rlive.size()));
	if (!ext_round(who, req, o, res)) return 0;
	for (size_t k = 0; k < who.size(); ++k)
		if (res[k].aa_len == who[k]->as1_win) who[k]->nt_len = res[k].nt_len, who[k]->aa_len = res[k].aa_len;
	for (size_t k = 0; k < rwho.size(); ++k) {
		const nsh_ext_result_t &x = res[who.size() + k];
		if (x.aa_len == rwho[k]->aa_win) rwho[k]->rnt_len = x.nt_len, rwho[k]->raa_len = x.aa_len;
	}
	clk.lap("extension retries", (int64_t)(who.size() + rwho.size()));
	// MP_AS_GAP (the DPs between consecutive trusted anchors, mp_align_gap, align.c:482-540) and
	// MP_AS_RIGHT_APPLY (the DP over the right extension) as one global-DP batch; a candidate's
	// tasks are its gap steps, then its right DP
	TaskSet ts;
	ts.first.assign(n_live + 1, 0);
	nsh_parallel_for((int64_t)n_live, nth, [&](int64_t k) {   // tasks per candidate
		const Cand *c = live[k];
		int64_t m = c->right;
		for (int32_t i = c->i0; i < c->r->cnt; ++i) m += (c->a[i] & kTrusted) != 0;
		ts.first[k + 1] = m;
	});
	for (size_t k = 0; k < n_live; ++k) ts.first[k + 1] += ts.first[k];
	ts.t.resize((size_t)ts.first[n_live]);
	nsh_parallel_for((int64_t)n_live, nth, [&](int64_t k) {
		Cand *c = live[k];
		nsh_align_reg_t *r = c->r;
		r->vs = c->vs1 - c->nt_len;
		r->qs = c->as1 - c->aa_len;
		c->ne0 = (int32_t)(r->vs - c->vs0);
		c->ae0 = r->qs;
		int64_t w = ts.first[k];
		for (int32_t i = c->i0; i < r->cnt; ++i) {
			if (!(c->a[i] & kTrusted)) continue;
			const int32_t ne1 = (int32_t)anchor_x(c->a[i]) + 1, ae1 = anchor_q(c->a[i]) + 1;
			Task &t = ts.t[w++];
			t.c = c, t.nt_off = c->ne0 + c->vs0 - c->as, t.nlen = ne1 - c->ne0, t.aa_off = c->ae0, t.alen = ae1 - c->ae0;
			t.score = 0, t.dp = false;
			c->i0 = i, c->ne0 = ne1, c->ae0 = ae1;
		}
		r->ve = c->ne0 + c->vs0, r->qe = c->ae0;   // the last trusted anchor's end: ve1, qe1
		if (c->right) {
			Task &t = ts.t[w++];
			t.c = c, t.nt_off = r->ve - c->as, t.nlen = c->rnt_len, t.aa_off = r->qe, t.alen = c->raa_len, t.score = 0, t.dp = false;
		}
	});
	clk.lap("DP tasks", (int64_t)ts.t.size());
	if (!dp_round(ts, o, ao, nth)) return 0;
	merge_tasks(ts, live, nth);
	for (Cand *c : rlive) c->r->ve += c->rnt_len, c->r->qe += c->raa_len;
	clk.lap("gap and right DPs", (int64_t)ts.t.size());
	// MP_AS_FINISH: the CIGARs, one after another
	int64_t tot = 0;
	for (Cand *c : live) tot += (int64_t)c->cigar.size();
	uint32_t *out = (uint32_t*)std::malloc((size_t)std::max<int64_t>(tot, 1) * sizeof(uint32_t));
	if (!out) { nsh_set_error("nsh_align_stage: out of host memory"); return 0; }
	int64_t off = 0;
	for (Cand *c : live) {
		nsh_align_reg_t *r = c->r;
		r->status = 1, r->dp_score = c->score, r->win_as = c->as, r->win_ae = c->ae;
		r->cigar_off = off, r->n_cigar = (int32_t)c->cigar.size();
		if (!c->cigar.empty()) std::memcpy(out + off, c->cigar.data(), c->cigar.size() * sizeof(uint32_t));
		off += (int64_t)c->cigar.size();
	}
	*cigar_out = out;
	return 1;
}

extern "C" int64_t nsh_align_records(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		const char *const *ctg_name, const char *const *qname, const char *const *aa, const int32_t *qlen,
		const nsh_align_reg_t *r, const uint32_t *cigar, const nsh_align_opt_t *ao, uint32_t flag, int64_t id,
		int32_t hit_idx, char *out, int64_t size)
{
	if (!r || r->status == 0 || (r->vid >> 1) >= (uint32_t)n_ctg) return 0;
	ns_opt_t o;
	ns_opt_init(&o);
	o.go = ao->go, o.ge = ao->ge, o.fs = ao->fs, o.sc = ao->mat;
	std::vector<uint8_t> win((size_t)std::max<int64_t>(1, r->win_ae - r->win_as));
	if (nsh_strand_bases(nt4, ctg_off, ctg_len, r->vid, r->win_as, r->win_ae, win.data()) != r->win_ae - r->win_as) {
		nsh_set_error("nsh_align_records: window outside its contig");
		return -1;
	}
	nsh_aln_t a = {};
	a.qname = qname[r->query], a.aa = aa[r->query], a.qlen = qlen[r->query], a.qs = r->qs, a.qe = r->qe;
	a.ctg_name = ctg_name[r->vid >> 1], a.ctg_len = ctg_len[r->vid >> 1], a.rev = r->vid & 1, a.vs = r->vs, a.ve = r->ve;
	a.dp_score = r->dp_score, a.profile_score = INT32_MIN, a.chn_sc = r->chn_sc, a.chn_sc_ungap = r->chn_sc_ungap;
	a.cnt = r->cnt, a.cigar = cigar + r->cigar_off, a.n_cigar = r->n_cigar;
	nsh_aln_stats_t st;
	std::vector<nsh_feat_t> feat(256);
	int32_t nf = nsh_aln_stats(&a, win.data(), r->win_as, r->win_ae, &o, nullptr, &st, feat.data(), (int32_t)feat.size());
	if (nf < 0) {   // more exons than room: once more with enough
		feat.resize(4096);
		nf = nsh_aln_stats(&a, win.data(), r->win_as, r->win_ae, &o, nullptr, &st, feat.data(), (int32_t)feat.size());
		if (nf < 0) return -1;
	}
	// mp_set_rank_score (hit.c:383-399): the aligned region ranks by its peak DP score
	a.rank_sc = nf > 0? st.dp_max : 0;
	return nsh_write_output(out, size, &a, nf > 0? &st : nullptr, feat.data(), win.data(), r->win_as, flag, 0, nullptr,
			id, -1, hit_idx);
}
