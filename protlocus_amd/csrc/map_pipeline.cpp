// The map path for a batch of proteins on one GPU (nsh_map_batch): what the reference's
// worker_pipeline does per mini-batch with its default options (map.c:1091-1180): mp_map_coarse
// (map.c:547-641), mp_map_finish (map.c:643-870) and the output step (map.c:1143-1166).
//
// The reference runs each query through every stage on one CPU thread. Here each stage runs for
// the whole batch at once, on the device where the work is data-parallel:
//   coarse   index -> anchors -> pre-chain -> chain of every query   (nsh_index_chain, device)
//            regions from the chain blocks, sort, parents, selection   (host, per query)
//   refine   the query sketches                                        (host)
//            window k-mers shared with the query, bucket pairs, chains  (refine.hip, device)
//   align    extensions and gap DPs of every candidate                 (nsh_align_stage, device)
//            statistics, ranking, selection and the PAF/GFF3 records   (host threads)
// Every step keeps the reference's arithmetic and order, so the records are the reference's byte
// for byte; where its results depend on the order a non-stable sort leaves equal keys in (the
// radix sort of ksort.h), that sort is restated step for step.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <memory>
#include <vector>
#include "nsh_internal.h"
#include "prep.h"

namespace {

struct P128 { uint64_t x, y; };

// radix_sort_mp128x (ksort.h:110-162, keyed on x): MSD, 8 bits a pass, buckets permuted in place by
// cycles, insertion sort for ranges of at most 64. Not stable: the order it leaves equal keys in is
// part of the result (which of two equally scored regions comes first), so the steps are kept.
void rs_insert(P128 *beg, P128 *end)
{
	for (P128 *i = beg + 1; i < end; ++i)
		if (i->x < (i - 1)->x) {
			P128 tmp = *i, *j = i;
			for (; j > beg && tmp.x < (j - 1)->x; --j) *j = *(j - 1);
			*j = tmp;
		}
}

void rs_sort(P128 *beg, P128 *end, int n_bits, int s)
{
	struct Bucket { P128 *b, *e; };
	const int size = 1 << n_bits, m = size - 1;
	Bucket b[256], *be = b + size;
	for (Bucket *k = b; k != be; ++k) k->b = k->e = beg;
	for (P128 *i = beg; i != end; ++i) ++b[i->x >> s & m].e;
	for (Bucket *k = b + 1; k != be; ++k) k->e += (k - 1)->e - beg, k->b = (k - 1)->e;
	for (Bucket *k = b; k != be;) {
		if (k->b != k->e) {
			Bucket *l = b + (k->b->x >> s & m);
			if (l != k) {
				P128 tmp = *k->b, swap;
				do {
					swap = tmp, tmp = *l->b, *l->b++ = swap;
					l = b + (tmp.x >> s & m);
				} while (l != k);
				*k->b++ = tmp;
			} else ++k->b;
		} else ++k;
	}
	b->b = beg;
	for (Bucket *k = b + 1; k != be; ++k) k->b = (k - 1)->e;
	if (s) {
		s = s > n_bits? s - n_bits : 0;
		for (Bucket *k = b; k != be; ++k)
			if (k->e - k->b > 64) rs_sort(k->b, k->e, n_bits, s);
			else if (k->e - k->b > 1) rs_insert(k->b, k->e);
	}
}

void radix_sort128(P128 *beg, P128 *end)
{
	if (end - beg <= 64) rs_insert(beg, end);
	else rs_sort(beg, end, 8, 56);
}

// one region of a query (the fields of mp_reg1_t the map path reads)
struct Reg {
	int32_t id = 0, parent = 0;
	uint32_t vid = 0;
	int32_t qs = 0, qe = 0, cnt = 0, chn_sc = 0, chn_sc_ungap = 0, rank_sc = 0;
	int64_t vs = 0, ve = 0, off = 0;   // off: the region's first anchor among the query's coarse anchors
	std::vector<uint64_t> a;           // refined anchors: x relative to vs << 32 | query position
	bool p = false;                    // aligned (mp_reg1_t.p != NULL)
	bool audited = false;              // rejected by the prescreen but kept by its audit
	int32_t n_exon = 0, slot = -1;     // slot: the candidate's index in the align stage
};

inline int32_t reg_score(const Reg &r) { return r.p? r.rank_sc : r.chn_sc; }   // mp_reg_score, hit.c:330-333

// A query's regions are an index array into its Reg pool, so the in-place compactions below move
// indices exactly as the reference moves mp_reg1_t structs.
using Regs = std::vector<int32_t>;

// mp_sort_reg without the rerank tie-break (hit.c:426-472): regions with cnt 0 are squeezed out,
// the rest ordered by descending score (mp_reg1_t.hash is 0 on this path)
void sort_reg(std::vector<Reg> &pool, Regs &r)
{
	const int32_t n = (int32_t)r.size();
	if (n <= 1) return;
	std::vector<P128> aux;
	aux.reserve(n);
	for (int32_t i = 0; i < n; ++i) {
		const Reg &x = pool[r[i]];
		if (x.cnt > 0) aux.push_back({ (uint64_t)(uint32_t)(x.p? reg_score(x) : x.chn_sc) << 32, (uint64_t)i });
	}
	radix_sort128(aux.data(), aux.data() + aux.size());
	Regs t(aux.size());
	for (size_t i = 0; i < aux.size(); ++i) t[aux.size() - 1 - i] = r[(uint32_t)aux[i].y];
	r.swap(t);
}

// mp_set_parent (hit.c:474-534) with hard_mask_level 0: ids are positions; a region whose query
// interval is mostly covered by an earlier primary becomes that primary's secondary
void set_parent(std::vector<Reg> &pool, const Regs &r, float mask_level, int32_t mask_len)
{
	const int32_t n = (int32_t)r.size();
	if (n <= 0) return;
	for (int32_t i = 0; i < n; ++i) pool[r[i]].id = i;
	std::vector<int32_t> w(n);
	std::vector<uint64_t> cov;
	w[0] = 0, pool[r[0]].parent = 0;
	int32_t k = 1;
	for (int32_t i = 1; i < n; ++i) {
		Reg &ri = pool[r[i]];
		const int32_t si = ri.qs, ei = ri.qe;
		int32_t uncov_len = 0, j;
		cov.clear();
		for (j = 0; j < k; ++j) {
			const Reg &rp = pool[r[w[j]]];
			int32_t sj = rp.qs, ej = rp.qe;
			if (ej <= si || sj >= ei) continue;
			if (sj < si) sj = si;
			if (ej > ei) ej = ei;
			cov.push_back((uint64_t)(uint32_t)sj << 32 | (uint32_t)ej);
		}
		if (!cov.empty()) {   // the query length not covered by overlapping primaries
			std::sort(cov.begin(), cov.end());
			int32_t x = si;
			for (uint64_t c : cov) {
				if ((int32_t)(c >> 32) > x) uncov_len += (int32_t)(c >> 32) - x;
				x = (int32_t)c > x? (int32_t)c : x;
			}
			if (ei > x) uncov_len += ei - x;
			for (j = 0; j < k; ++j) {
				const Reg &rp = pool[r[w[j]]];
				const int32_t sj = rp.qs, ej = rp.qe;
				if (ej <= si || sj >= ei) continue;
				const int32_t mn = ej - sj < ei - si? ej - sj : ei - si, mx = ej - sj > ei - si? ej - sj : ei - si;
				const int32_t ol = si < sj? (ei < sj? 0 : ei < ej? ei - sj : ej - sj) : (ej < si? 0 : ej < ei? ej - si : ei - si);
				if ((float)ol / mn - (float)uncov_len / mx > mask_level && uncov_len <= mask_len) {
					ri.parent = rp.parent;
					break;
				}
			}
		} else j = k;
		if (j == k) w[k++] = i, ri.parent = i;
	}
}

// mp_sync_regs (hit.c:536-557)
void sync_regs(std::vector<Reg> &pool, const Regs &r)
{
	const int32_t n = (int32_t)r.size();
	if (n <= 0) return;
	int32_t max_id = -1;
	for (int32_t i = 0; i < n; ++i) max_id = std::max(max_id, pool[r[i]].id);
	std::vector<int32_t> tmp((size_t)max_id + 1, -1);
	for (int32_t i = 0; i < n; ++i)
		if (pool[r[i]].id >= 0) tmp[pool[r[i]].id] = i;
	for (int32_t i = 0; i < n; ++i) {
		Reg &x = pool[r[i]];
		x.id = i;
		if (x.parent == -2) x.parent = i;
		else if (x.parent >= 0 && x.parent <= max_id && tmp[x.parent] >= 0) x.parent = tmp[x.parent];
		else x.parent = -1;
	}
}

// mp_select_sub (hit.c:559-593): secondaries kept when close enough to their primary, at most
// best_n of them. The compaction is in place and parents are looked up by position while it runs,
// as in the reference.
void select_sub(std::vector<Reg> &pool, Regs &r, float pri_ratio, int32_t min_diff, int32_t best_n)
{
	const int32_t n = (int32_t)r.size();
	if (!(pri_ratio > 0.0f) || n <= 0) return;
	int32_t k = 0, n_2nd = 0, ungap = -1;
	for (int32_t i = 0; i < n; ++i) ungap = std::max(ungap, pool[r[i]].chn_sc_ungap);
	auto same = [&](const Reg &a, const Reg &b) {
		return a.qs == b.qs && a.qe == b.qe && a.vid == b.vid && a.vs == b.vs && a.ve == b.ve;
	};
	for (int32_t i = 0; i < n; ++i) {
		const Reg &ri = pool[r[i]];
		const int32_t p = ri.parent;
		if (p < 0) { r[k++] = r[i]; continue; }
		const Reg &rp = pool[r[p]];
		const int32_t sci = reg_score(ri), scp = reg_score(rp);
		if (p == i) {
			r[k++] = r[i];
		} else if ((sci >= scp * pri_ratio || sci + min_diff >= scp) && n_2nd < best_n) {
			if (!same(ri, rp)) r[k++] = r[i], ++n_2nd;
		} else if (!ri.p && !rp.p && ungap > 0 && ri.chn_sc_ungap >= ungap * pri_ratio && n_2nd < best_n) {
			if (!same(ri, rp)) r[k++] = r[i], ++n_2nd;
		}
	}
	if (k != n) {
		r.resize(k);
		sync_regs(pool, r);
	}
}

// mp_cal_max_ext (hit.c:744-781): how far each region may be extended, left << 32 | right. With
// by_pos the regions are ordered by their place in the genome, otherwise by their first coarse
// anchor's block (a), as mp_map_finish calls it before and after the refinement
std::vector<uint64_t> cal_max_ext(const std::vector<Reg> &pool, const Regs &r, bool by_pos, const int64_t *ctg_off,
		const int64_t *ctg_len, const uint64_t *a, int32_t min_ext, int32_t max_ext, int32_t iso_ext)
{
	const int32_t n = (int32_t)r.size();
	std::vector<uint64_t> ext(n);
	if (n <= 0) return ext;
	std::vector<P128> b(n);
	for (int32_t i = 0; i < n; ++i) {
		const Reg &x = pool[r[i]];
		b[i].x = by_pos? (uint64_t)(x.vs + ctg_off[x.vid >> 1] + (x.vid & 1? ctg_len[x.vid >> 1] : 0)) : a[x.off] >> 32;
		b[i].y = (uint64_t)i;
	}
	radix_sort128(b.data(), b.data() + n);
	for (int32_t i = 0; i < n; ++i) {
		int32_t left = iso_ext, right = iso_ext;
		const int32_t j = (int32_t)b[i].y;
		const Reg &x = pool[r[j]];
		if (i > 0) {
			const Reg &q = pool[r[(int32_t)b[i - 1].y]];
			if (q.vid == x.vid && q.qe >= x.qs) {
				left = (int32_t)(x.vs - q.ve < max_ext? x.vs - q.ve : max_ext);
				left = left > min_ext? left : min_ext;
			}
		}
		if (i < n - 1) {
			const Reg &q = pool[r[(int32_t)b[i + 1].y]];
			if (q.vid == x.vid && x.qe >= q.qs) {
				right = (int32_t)(q.vs - x.ve < max_ext? q.vs - x.ve : max_ext);
				right = right > min_ext? right : min_ext;
			}
		}
		ext[j] = (uint64_t)(uint32_t)left << 32 | (uint32_t)right;
	}
	return ext;
}

// mp_idx_block2pos (index.c:31-47): the strand (contig << 1 | rev) a block belongs to
int32_t block2pos(const std::vector<uint32_t> &bo, uint32_t b)
{
	const int32_t n = (int32_t)bo.size() - 1;
	if (b >= bo[n]) return -1;
	int32_t st = 0, en = n - 1;
	while (st <= en) {
		const int32_t mid = (st + en) / 2;
		if (bo[mid] <= b && b < bo[mid + 1]) return mid;
		else if (b < bo[mid]) en = mid - 1;
		else st = mid + 1;
	}
	return -2;
}

// mp_reg_gen_from_block (hit.c:265-309) with mp_cal_chn_sc_ungap_approx (hit.c:6-16): one region
// per chain; a chain over two strands keeps the side with more anchors
bool regs_from_chains(const std::vector<uint32_t> &bo, int32_t bbit, int32_t kmer, int32_t n_u, const uint64_t *u,
		const uint64_t *a, std::vector<Reg> &pool, Regs &r)
{
	int64_t k = 0;
	for (int32_t i = 0; i < n_u; ++i) {
		const uint32_t n = (uint32_t)u[i];
		int64_t is = k, ie = k + n - 1;
		const int32_t ts = block2pos(bo, (uint32_t)(a[is] >> 32)), te = block2pos(bo, (uint32_t)(a[ie] >> 32));
		if (ts < 0 || te < 0 || ts > te) return false;
		Reg x;
		x.off = k, x.cnt = (int32_t)n;
		if (ts == te) {
			x.vid = (uint32_t)ts;
		} else {
			int64_t j = k;
			for (; j < k + n; ++j)
				if (a[j] >> 32 >= bo[ts + 1]) break;
			const int64_t js = j;
			for (j = k + n - 1; j >= js; --j)
				if (a[j] >> 32 < bo[te]) break;
			const int64_t je = j + 1;
			if (js - k > k + n - je) x.vid = (uint32_t)ts, ie = js - 1;
			else x.vid = (uint32_t)te, is = je;
		}
		x.vs = (int64_t)((a[is] >> 32) - bo[x.vid]) << bbit;
		x.ve = (int64_t)((a[ie] >> 32) - bo[x.vid] + 1) << bbit;
		x.qs = (int32_t)(uint32_t)a[is];
		x.qe = (int32_t)(uint32_t)a[ie];
		x.chn_sc = ts == te? (int32_t)(u[i] >> 32) : (int32_t)(uint32_t)((double)(u[i] >> 32) * (double)(ie - is + 1) / n + .499);
		int32_t s = kmer;
		for (int32_t t = 1; t < x.cnt; ++t) {
			const uint64_t a0 = a[x.off + t - 1], a1 = a[x.off + t];
			const int32_t dq = (int32_t)a1 - (int32_t)a0;
			s += dq < kmer? dq : kmer;
			if (a1 >> 32 == a0 >> 32) s += 2;   // MP_BLOCK_BONUS
		}
		x.chn_sc_ungap = s;
		r.push_back((int32_t)pool.size());
		pool.push_back(std::move(x));
		k += n;
	}
	return true;
}

// mp_cal_chn_sc_ungap (hit.c:18-32) of refined anchors
int32_t chn_sc_ungap(const std::vector<uint64_t> &a, int32_t kmer)
{
	int32_t x = kmer;
	for (size_t i = 1; i < a.size(); ++i) {
		const int32_t dq = (int32_t)a[i] - (int32_t)a[i - 1], dr3 = (int32_t)((a[i] >> 32) - (a[i - 1] >> 32));
		const int32_t dr = dr3 / 3, q = dr3 - dr * 3, dg = dq < dr? dq : dr;
		if (dq >= dr && q != 0) --x;
		else x += dg < kmer? dg : kmer;
	}
	return x;
}

// what the align stage leaves for the output of one aligned candidate
struct Out {
	nsh_aln_stats_t st;
	std::vector<nsh_feat_t> feat;
};

}   // namespace

struct nsh_mapper_s {
	int32_t n_ctg = 0, kmer = 0, bbit = 0, min_aa_len = 0;
	std::vector<std::string> name;
	std::vector<const char*> name_p;
	std::vector<uint8_t> nt;                 // base codes 0..4: the contigs one after another, then their reverse complements
	std::vector<int64_t> off, len;
	std::vector<uint32_t> bo;                // first block of each strand, 2 n_ctg + 1 entries
	int64_t total = 0;                       // bases
	nsh_index_t *ix = nullptr;
	uint8_t *d_g = nullptr;                  // HBM: the contigs, then their reverse complements
	nsh_genome_view view;                    // nt (both orientations) and d_g
};

extern "C" void nsh_map_opt_default(nsh_map_opt_t *o)
{
	std::memset(o, 0, sizeof(*o));
	nsh_align_opt_default(&o->dp);               // options.c:42-138
	o->max_occ = 20000, o->max_gap = 1000, o->bw = o->dp.max_intron;
	o->max_chn_max_skip = 25, o->max_chn_iter = 1000000, o->min_chn_cnt = 3, o->min_chn_sc = 0;
	o->chn_coef_log = 0.75f;
	o->max_ava = 1000, o->best_n = 30, o->out_n = 1000, o->mask_len = INT32_MAX;
	o->pri_ratio = 0.7f, o->mask_level = 0.5f, o->out_sim = 0.99f, o->out_cov = 0.1f;
	o->flag = 0, o->gff_delim = -1, o->gff_prefix = nullptr;
	o->n_threads = 16;
	o->ps_audit = 0.02f;
}

extern "C" nsh_mapper_t *nsh_mapper_init(int32_t n_ctg, const char *const *ctg_name, const uint8_t *nt4,
		const int64_t *ctg_len, int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
{
	if (n_ctg <= 0 || !ctg_name || !nt4 || !ctg_len || kmer < 1 || kmer > 7 || bbit < 0 || bbit > 30) {
		nsh_set_error("nsh_mapper_init: bad arguments");
		return nullptr;
	}
	nsh_mapper_t *m = new nsh_mapper_t;
	m->n_ctg = n_ctg, m->kmer = kmer, m->bbit = bbit, m->min_aa_len = min_aa_len;
	int64_t tot = 0;
	for (int32_t c = 0; c < n_ctg; ++c) {
		if (ctg_len[c] <= 0 || !ctg_name[c]) { nsh_set_error("nsh_mapper_init: contig %d is empty", c); delete m; return nullptr; }
		m->name.emplace_back(ctg_name[c]);
		m->off.push_back(tot), m->len.push_back(ctg_len[c]);
		tot += ctg_len[c];
	}
	for (auto &s : m->name) m->name_p.push_back(s.c_str());
	m->nt.resize((size_t)(2 * tot));
	std::memcpy(m->nt.data(), nt4, (size_t)tot);
	nsh_parallel_for(n_ctg, 16, [&](int64_t c) {
		nsh_strand_bases(nt4, m->off.data(), m->len.data(), (uint32_t)c << 1 | 1, 0, m->len[c], m->nt.data() + tot + m->off[c]);
	});
	m->total = tot;
	if (!nsh_refine::oriented_genome(m->nt.data(), m->off.data(), m->len.data(), n_ctg, tot, &m->d_g)) { delete m; return nullptr; }
	m->ix = nsh_index_build(n_ctg, m->nt.data(), m->off.data(), m->len.data(), min_aa_len, kmer, mod_bit, bbit);
	if (!m->ix) { nsh_mapper_destroy(m); return nullptr; }
	m->view.n_ctg = n_ctg, m->view.off = m->off.data(), m->view.len = m->len.data(), m->view.total = tot;
	m->view.h_g = m->nt.data(), m->view.d_g = m->d_g;
	m->bo.resize(2 * (size_t)n_ctg + 1);
	if (!nsh_index_copy(m->ix, nullptr, nullptr, m->bo.data())) { nsh_mapper_destroy(m); return nullptr; }
	return m;
}

extern "C" void nsh_mapper_destroy(nsh_mapper_t *m)
{
	if (!m) return;
	if (m->ix) nsh_index_destroy(m->ix);
	if (m->d_g) (void)hipFree(m->d_g);
	delete m;
}

extern "C" int nsh_map_batch(nsh_mapper_t *m, int32_t n_q, const char *const *qname, const char *const *seq,
		const int32_t *qlen, const nsh_map_opt_t *o, int64_t *id, char **out, int64_t *out_len, int64_t *stats)
{
	if (!m || n_q < 0 || (n_q > 0 && (!qname || !seq || !qlen)) || !o || !id || !out || !out_len) {
		nsh_set_error("nsh_map_batch: bad arguments");
		return 0;
	}
	*out = nullptr, *out_len = 0;
	const int32_t nth = o->n_threads > 0? o->n_threads : 1;
	const int32_t kmer2 = o->dp.kmer2;
	int64_t st[24] = { 0 };
	const auto t_start = std::chrono::steady_clock::now();
	auto lap = [&](int k) {   // microseconds since the call began, at the end of stage k
		st[8 + k] = (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t_start).count();
	};
	if (n_q == 0) {
		*out = (char*)std::malloc(1);
		if (stats) std::memcpy(stats, st, sizeof(st));
		return *out != nullptr;
	}

	// ---- coarse: index -> anchors -> pre-chain -> chain, all queries on the device (map.c:583-633)
	std::vector<char> qs_all;
	std::vector<int64_t> q_off(n_q);
	for (int32_t q = 0; q < n_q; ++q) {
		q_off[q] = (int64_t)qs_all.size();
		qs_all.insert(qs_all.end(), seq[q], seq[q] + qlen[q]);
	}
	qs_all.push_back(0);
	nsh_chain_opt_t co = { o->dp.max_intron, o->max_gap, o->bw, o->max_chn_max_skip, o->max_chn_iter, o->min_chn_cnt,
		o->min_chn_sc, o->chn_coef_log, 1, 1 };
	std::vector<int64_t> u_off(n_q + 1), c_off(n_q + 1);
	// the chains and their anchors: room for a generous bound, left uninitialised (a zero-filled
	// vector of that bound cost about 0.1 s per 20,000 queries; only the filled part is touched)
	std::unique_ptr<uint64_t[]> cu_buf, ca_buf;
	uint64_t *cu = nullptr, *ca = nullptr;
	int64_t cap_u = 64 * (int64_t)n_q + 1024, cap_a = 4096 * (int64_t)n_q + 65536;
	for (;;) {
		cu_buf.reset(new uint64_t[(size_t)cap_u]), ca_buf.reset(new uint64_t[(size_t)cap_a]);
		cu = cu_buf.get(), ca = ca_buf.get();
		const int64_t nu = nsh_index_chain(m->ix, n_q, qs_all.data(), q_off.data(), qlen, o->max_occ, &co, u_off.data(),
				cu, cap_u, c_off.data(), ca, cap_a);
		if (nu < 0) return 0;
		if (nu <= cap_u && c_off[n_q] <= cap_a) break;
		cap_u = std::max(cap_u, nu), cap_a = std::max(cap_a, c_off[n_q]);
	}
	st[0] = u_off[n_q];
	lap(0);

	std::vector<std::vector<Reg>> pool(n_q);
	std::vector<Regs> regs(n_q);
	const int32_t sel_n = o->best_n;
	const float sel_ratio = o->pri_ratio;
	std::atomic<bool> bad(false);
	nsh_parallel_for(n_q, nth, [&](int64_t q) {   // mp_reg_gen_from_block, sort, parents, selection (map.c:634-639)
		if (!regs_from_chains(m->bo, m->bbit, m->kmer, (int32_t)(u_off[q + 1] - u_off[q]), cu + u_off[q],
				ca + c_off[q], pool[q], regs[q])) { bad = true; return; }
		sort_reg(pool[q], regs[q]);
		set_parent(pool[q], regs[q], o->mask_level, o->mask_len);
		select_sub(pool[q], regs[q], sel_ratio * sel_ratio, m->kmer * 2, sel_n);
	});
	if (bad) { nsh_set_error("nsh_map_batch: a chain block is outside the index"); return 0; }
	lap(1);

	// ---- the learned prescreen between coarse chaining and refinement (where the reference's
	// --gpu-prefilter sits, map.c:1114-1117): every candidate's 64 features (the window-sequence
	// features of the device kernel, slots 46-59 replaced by the chain's own), the network, the cut
	// and a random audit of what it rejects
	std::vector<std::pair<int32_t, int32_t>> cands;   // (query, pool index) of every candidate offered
	if (o->prescreen || (o->cand_feat && o->cand_cap > 0)) {
		for (int32_t q = 0; q < n_q; ++q)
			for (int32_t i : regs[q]) cands.push_back({ q, i });
		const int32_t nc = (int32_t)cands.size();
		std::vector<int64_t> g_off(nc), a_off(nc);
		std::vector<int32_t> wl(nc), al(nc);
		for (int32_t c = 0; c < nc; ++c) {
			const int32_t q = cands[c].first;
			const Reg &r = pool[q][cands[c].second];
			const int64_t clen = m->len[r.vid >> 1], vs = std::min<int64_t>(std::max<int64_t>(r.vs, 0), clen);
			const int64_t ve = std::min<int64_t>(std::max<int64_t>(r.ve, vs), clen);
			g_off[c] = m->view.at(r.vid, vs);
			wl[c] = (int32_t)std::min<int64_t>(ve - vs, nsh_feat::kMaxWindow);
			a_off[c] = q_off[q], al[c] = std::max(1, std::min(qlen[q], nsh_feat::kMaxProtein));
		}
		std::vector<float> feat((size_t)std::max(nc, 1) * 64);
		if (nc > 0 && !nsh_feat::from_genome(m->d_g, nc, g_off.data(), wl.data(), qs_all.data(), a_off.data(), al.data(),
				(int64_t)qs_all.size(), feat.data()))
			return 0;
		std::vector<int64_t> cfirst(n_q + 1, 0);   // first candidate of each query (candidates are in query order)
		for (int32_t q = 0; q < n_q; ++q) cfirst[q + 1] = cfirst[q] + (int64_t)regs[q].size();
		nsh_parallel_for(n_q, nth, [&](int64_t q) {   // the chain's own features
			const Regs &R = regs[q];
			int32_t best = 1;
			for (int32_t i : R) best = std::max(best, pool[q][i].chn_sc);
			const double L = std::max(qlen[q], 1), n = (double)R.size();
			const int64_t c0 = cfirst[q];
			for (size_t k = 0; k < R.size(); ++k) {
				const Reg &r = pool[q][R[k]];
				const Reg &par = pool[q][R[std::max(0, std::min<int32_t>(r.parent, (int32_t)R.size() - 1))]];
				const double span = (double)(r.ve - r.vs);
				float *f = feat.data() + (c0 + (int64_t)k) * 64;
				f[46] = (float)(r.chn_sc / L), f[47] = (float)((double)r.chn_sc / best);
				f[48] = (float)(r.chn_sc_ungap / L), f[49] = (float)(r.cnt / L);
				f[50] = (float)((r.qe - r.qs) / L), f[51] = (float)(r.qs / L), f[52] = (float)((L - r.qe) / L);
				f[53] = (float)std::log2(1.0 + span), f[54] = (float)(span / (3.0 * (r.qe - r.qs) + 1.0));
				f[55] = r.parent == (int32_t)k? 1.0f : 0.0f, f[56] = (float)(k / n), f[57] = (float)std::log2(1.0 + n);
				f[58] = (float)((double)r.chn_sc / std::max(par.chn_sc, 1)), f[59] = (float)((best - r.chn_sc) / L);
			}
		});
		st[6] = nc;
		if (o->cand_feat) std::memcpy(o->cand_feat, feat.data(), (size_t)std::min<int64_t>(nc, o->cand_cap) * 64 * 4);
		if (o->prescreen && nc > 0) {
			std::vector<float> logit(nc);
			if (!nsh_mlp_infer(o->prescreen, feat.data(), nc, 0, logit.data())) return 0;
			std::vector<char> keep(nc);
			for (int32_t c = 0; c < nc; ++c) {
				keep[c] = logit[c] >= o->ps_tau;
				if (!keep[c]) {   // the audit: a fixed fraction of the rejected, drawn by a hash of the seed and the candidate
					uint64_t h = o->ps_seed ^ ((uint64_t)(uint32_t)cands[c].first << 32 | (uint32_t)cands[c].second);
					h ^= h >> 33, h *= 0xff51afd7ed558ccdULL, h ^= h >> 33, h *= 0xc4ceb9fe1a85ec53ULL, h ^= h >> 33;
					if ((double)(h >> 11) * (1.0 / 9007199254740992.0) < o->ps_audit) keep[c] = 2, ++st[14];
					else ++st[7];
				}
			}
			for (int32_t q = 0, c = 0; q < n_q; ++q) {
				Regs kept;
				for (int32_t i : regs[q]) {
					if (keep[c]) kept.push_back(i);
					if (keep[c] == 2) pool[q][i].audited = true;
					++c;
				}
				regs[q].swap(kept);
			}
		}
	}
	if (!cands.empty()) st[16] = (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t_start).count();

	// ---- refine: every region's window against its query (map.c:141-258, 666-684)
	nsh_step_clock hclk("nsh_map_batch");
	hclk.lap("  (region selection done)");
	struct RefJob { int32_t q, r; int64_t as; };
	std::vector<RefJob> jobs;
	std::vector<std::vector<uint64_t>> ext0(n_q);
	nsh_parallel_for(n_q, nth, [&](int64_t q) {
		ext0[q] = cal_max_ext(pool[q], regs[q], false, nullptr, nullptr, ca + c_off[q], 100, o->dp.max_ext,
				o->dp.max_ext);
	});
	for (int32_t q = 0; q < n_q; ++q)
		for (int32_t i = 0; i < (int32_t)regs[q].size(); ++i) jobs.push_back({ q, i, 0 });
	st[1] = (int64_t)jobs.size();
	hclk.lap("  extension limits", (int64_t)jobs.size());
	// the windows (mp_refine_reg, map.c:152-156): slices of the oriented genome in HBM
	std::vector<nsh_refine::Job> rjobs(jobs.size());
	nsh_parallel_for(((int64_t)jobs.size() + 4095) / 4096, nth, [&](int64_t b) {
	for (size_t t = (size_t)b * 4096; t < std::min(jobs.size(), (size_t)(b + 1) * 4096); ++t) {
		RefJob &jb = jobs[t];
		const Reg &r = pool[jb.q][regs[jb.q][jb.r]];
		const int32_t extl = (int32_t)(ext0[jb.q][jb.r] >> 32), extr = (int32_t)ext0[jb.q][jb.r];
		const int64_t clen = m->len[r.vid >> 1];
		const int64_t as = r.vs > extl? r.vs - extl : 0, ae = r.ve + extr < clen? r.ve + extr : clen;
		jb.as = as;
		rjobs[t].g_off = (r.vid & 1? m->total : 0) + m->off[r.vid >> 1] + as;
		rjobs[t].len = (int32_t)std::max<int64_t>(0, ae - as);
		rjobs[t].q = jb.q;
	}
	});
	hclk.lap("  windows", (int64_t)jobs.size());
	lap(2);
	// window k-mers shared with the query, bucket pairs and the window chains of every region (map.c:153-203)
	std::vector<uint64_t> best_u, best_a;
	std::vector<int64_t> best_off;
	const nsh_chain::Params rcp = { o->dp.max_intron, o->max_gap, o->bw, o->max_chn_max_skip, o->max_chn_iter,
		o->min_chn_cnt, o->min_chn_sc, o->chn_coef_log, 1, kmer2, 0 };
	if (!nsh_refine::run(m->d_g, rjobs, qs_all.data(), q_off.data(), qlen, n_q, (int64_t)qs_all.size(), kmer2,
			m->min_aa_len, o->max_ava, rcp, best_u, best_off, best_a,
			&st[2]))
		return 0;
	hclk.lap("  refine on the device", (int64_t)jobs.size());
	nsh_parallel_for((int64_t)jobs.size(), nth, [&](int64_t t) {   // the best chain becomes the region (map.c:249-258)
		const RefJob &jb = jobs[t];
		Reg &r = pool[jb.q][regs[jb.q][jb.r]];
		const int32_t n_a = (int32_t)(uint32_t)best_u[t];
		if (n_a == 0) { r.cnt = 0, r.a.clear(); return; }
		const uint64_t *a = best_a.data() + best_off[t];
		r.chn_sc = (int32_t)(best_u[t] >> 32);
		r.rank_sc = 0;
		r.cnt = n_a;
		r.qs = (int32_t)((uint32_t)a[0] - (uint32_t)(kmer2 - 1));
		r.qe = (int32_t)((uint32_t)a[n_a - 1] + 1);
		r.vs = jb.as + (int64_t)(a[0] >> 32) + 1 - 3 * kmer2;
		r.ve = jb.as + (int64_t)(a[n_a - 1] >> 32) + 1;
		r.a.resize(n_a);
		for (int32_t i = 0; i < n_a; ++i)
			r.a[i] = (uint64_t)((int64_t)(a[i] >> 32) + jb.as - r.vs) << 32 | (a[i] & 0xffffffffULL);
		r.chn_sc_ungap = chn_sc_ungap(r.a, kmer2);
	});
	nsh_parallel_for(n_q, nth, [&](int64_t q) {   // map.c:685-698
		Regs kept;
		for (int32_t i : regs[q])
			if (pool[q][i].cnt > 0) kept.push_back(i);
		regs[q].swap(kept);
		sort_reg(pool[q], regs[q]);
		set_parent(pool[q], regs[q], o->mask_level, o->mask_len);
		select_sub(pool[q], regs[q], sel_ratio * sel_ratio, m->kmer * 2, sel_n);
	});

	hclk.lap("  refined regions", n_q);
	lap(3);
	// ---- align: every candidate of every query through the device stages (map.c:701-739)
	// candidates in query order: first candidate and first anchor of each query, then the fill
	std::vector<int64_t> cfirst(n_q + 1, 0), afirst(n_q + 1, 0);
	nsh_parallel_for(n_q, nth, [&](int64_t q) {
		int64_t na = 0;
		for (int32_t i : regs[q]) na += (int64_t)pool[q][i].a.size();
		cfirst[q + 1] = (int64_t)regs[q].size(), afirst[q + 1] = na;
	});
	for (int32_t q = 0; q < n_q; ++q) cfirst[q + 1] += cfirst[q], afirst[q + 1] += afirst[q];
	std::vector<nsh_align_reg_t> cand((size_t)cfirst[n_q]);
	std::unique_ptr<uint64_t[]> canch(new uint64_t[(size_t)afirst[n_q] + 1]);
	canch[afirst[n_q]] = 0;
	nsh_parallel_for(n_q, nth, [&](int64_t q) {
		const std::vector<uint64_t> ext = cal_max_ext(pool[q], regs[q], true, m->off.data(), m->len.data(), nullptr, 100,
				o->dp.max_intron / 2, o->dp.max_intron / 2);
		int64_t ao = afirst[q];
		for (size_t i = 0; i < regs[q].size(); ++i) {
			Reg &r = pool[q][regs[q][i]];
			nsh_align_reg_t &c = cand[(size_t)cfirst[q] + i];
			std::memset(&c, 0, sizeof(c));
			c.query = (int32_t)q, c.vid = r.vid, c.qs = r.qs, c.qe = r.qe, c.vs = r.vs, c.ve = r.ve, c.cnt = r.cnt;
			c.chn_sc = r.chn_sc, c.chn_sc_ungap = r.chn_sc_ungap, c.a_off = ao, c.ext = ext[i];
			std::copy(r.a.begin(), r.a.end(), canch.get() + ao);
			ao += (int64_t)r.a.size();
			r.slot = (int32_t)(cfirst[q] + (int64_t)i);
		}
	});
	st[3] = (int64_t)cand.size();
	uint32_t *cig = nullptr;
	std::vector<const char*> aa_p(seq, seq + n_q);
	hclk.lap("  align candidates", (int64_t)cand.size());
	if (!cand.empty() && !nsh_align_stage_view(m->view, aa_p.data(), qlen, (int32_t)cand.size(), cand.data(), canch.get(),
			&o->dp, &cig))
		return 0;
	hclk.lap("  align stage", (int64_t)cand.size());
	lap(4);
	ns_opt_t so;
	ns_opt_init(&so);
	so.go = o->dp.go, so.ge = o->dp.ge, so.fs = o->dp.fs, so.sc = o->dp.mat;
	std::vector<Out> res(cand.size());
	std::atomic<bool> fail(false);
	nsh_parallel_for((int64_t)cand.size(), nth, [&](int64_t k) {   // mp_align_finish: statistics, rank score
		const nsh_align_reg_t &c = cand[k];
		if (!c.status) return;
		const uint8_t *win = m->nt.data() + m->view.at(c.vid, c.win_as);
		nsh_aln_t al = {};
		al.qname = qname[c.query], al.aa = seq[c.query], al.qlen = qlen[c.query], al.qs = c.qs, al.qe = c.qe;
		al.ctg_name = m->name_p[c.vid >> 1], al.ctg_len = m->len[c.vid >> 1], al.rev = c.vid & 1;
		al.vs = c.vs, al.ve = c.ve, al.dp_score = c.dp_score, al.profile_score = INT32_MIN;
		al.cigar = cig + c.cigar_off, al.n_cigar = c.n_cigar;
		Out &r = res[k];
		r.feat.resize(64);
		int32_t nf = nsh_aln_stats(&al, win, c.win_as, c.win_ae, &so, nullptr, &r.st, r.feat.data(), 64);
		if (nf < 0) {
			r.feat.resize(8192);
			nf = nsh_aln_stats(&al, win, c.win_as, c.win_ae, &so, nullptr, &r.st, r.feat.data(), 8192);
			if (nf < 0) { fail = true; return; }
		}
		r.feat.resize(std::max(nf, 0));
		if (nf == 0) r.st.n_exon = -1;   // the CIGAR does not span the region: the reference drops the details
	});
	if (fail) { std::free(cig); nsh_set_error("nsh_map_batch: alignment statistics failed"); return 0; }
	hclk.lap("  alignment statistics", (int64_t)cand.size());

	// ---- after the DP: ranking and selection per query (map.c:840-866), then the records (map.c:1143-1166)
	std::vector<std::string> text(n_q);
	std::vector<int32_t> n_out(n_q, 0);
	nsh_parallel_for(n_q, nth, [&](int64_t q) {
		Regs kept;
		for (int32_t i : regs[q]) {
			Reg &r = pool[q][i];
			const nsh_align_reg_t &c = cand[r.slot];
			r.p = c.status && res[r.slot].st.n_exon >= 0;
			if (!r.p) continue;
			r.qs = c.qs, r.qe = c.qe, r.vs = c.vs, r.ve = c.ve;
			r.rank_sc = res[r.slot].st.dp_max;   // mp_set_rank_score: dp_max without --accuracy (hit.c:383-392)
			r.n_exon = res[r.slot].st.n_exon;
			kept.push_back(i);
		}
		regs[q].swap(kept);
		Regs &rq = regs[q];
		sort_reg(pool[q], rq);
		// mp_select_multi_exon without the strict mode (hit.c:606-624)
		if (rq.size() >= 2 && pool[q][rq[0]].n_exon == 1) {
			size_t i = 1;
			while (i < rq.size() && pool[q][rq[i]].n_exon < 2) ++i;
			if (i < rq.size() && reg_score(pool[q][rq[0]]) < reg_score(pool[q][rq[i]]) + o->dp.io) std::swap(rq[0], rq[i]);
		}
		set_parent(pool[q], rq, o->mask_level, o->mask_len);
		select_sub(pool[q], rq, sel_ratio, m->kmer * 2, sel_n);
	});
	hclk.lap("  ranking", n_q);
	// the records: ids run over the batch in query order, as the output step numbers them
	std::vector<int64_t> id0(n_q + 1, 0);
	std::vector<std::vector<int32_t>> emit(n_q);
	for (int32_t q = 0; q < n_q; ++q) {
		const Regs &rq = regs[q];
		const int32_t best = rq.empty()? -1 : reg_score(pool[q][rq[0]]);
		for (size_t j = 0; j < rq.size() && (int32_t)j < o->out_n; ++j) {
			const Reg &r = pool[q][rq[j]];
			const int32_t sc = reg_score(r);
			if (sc <= 0 || sc < (double)best * o->out_sim) continue;
			if (r.qe - r.qs < (double)qlen[q] * o->out_cov) continue;
			emit[q].push_back((int32_t)j);
		}
		id0[q + 1] = id0[q] + (int64_t)emit[q].size();
	}
	if (!cands.empty()) {   // which candidates became records: the labels, and what the audit caught
		std::vector<std::vector<char>> rec(n_q);
		for (int32_t q = 0; q < n_q; ++q) {
			rec[q].assign(pool[q].size(), 0);
			for (int32_t j : emit[q]) rec[q][regs[q][j]] = 1;
		}
		int64_t caught = 0;
		for (size_t c = 0; c < cands.size(); ++c) {
			const char r = rec[cands[c].first][cands[c].second];
			if (o->cand_label && (int64_t)c < o->cand_cap) o->cand_label[c] = (uint8_t)r;
			if (r && pool[cands[c].first][cands[c].second].audited) ++caught;
		}
		st[15] = caught;
	}
	nsh_parallel_for(n_q, nth, [&](int64_t q) {
		std::string &s = text[q];
		std::vector<char> buf(1 << 16);
		for (size_t e = 0; e < emit[q].size(); ++e) {
			const int32_t j = emit[q][e];
			const Reg &r = pool[q][regs[q][j]];
			const nsh_align_reg_t &c = cand[r.slot];
			const uint8_t *win = m->nt.data() + m->view.at(c.vid, c.win_as);
			nsh_aln_t al = {};
			al.qname = qname[q], al.aa = seq[q], al.qlen = qlen[q], al.qs = r.qs, al.qe = r.qe;
			al.ctg_name = m->name_p[r.vid >> 1], al.ctg_len = m->len[r.vid >> 1], al.rev = r.vid & 1;
			al.vs = r.vs, al.ve = r.ve, al.dp_score = c.dp_score, al.profile_score = INT32_MIN, al.rank_sc = r.rank_sc;
			al.chn_sc = r.chn_sc, al.chn_sc_ungap = r.chn_sc_ungap, al.cnt = r.cnt;
			al.cigar = cig + c.cigar_off, al.n_cigar = c.n_cigar;
			const Out &x = res[r.slot];
			for (;;) {
				const int64_t n = nsh_write_output(buf.data(), (int64_t)buf.size(), &al, &x.st, x.feat.data(), win,
						c.win_as, (int32_t)o->flag, 0, o->gff_prefix, *id + id0[q] + (int64_t)e + 1, o->gff_delim, j + 1);
				if (n < 0) { fail = true; return; }
				if (n < (int64_t)buf.size()) { s.append(buf.data(), (size_t)n); break; }
				buf.resize((size_t)n + 1);
			}
		}
		n_out[q] = (int32_t)emit[q].size();
	});
	std::free(cig);
	if (fail) { nsh_set_error("nsh_map_batch: writing a record failed"); return 0; }
	hclk.lap("  records", n_q);
	int64_t tot = 0;
	for (const auto &s : text) tot += (int64_t)s.size();
	char *o_buf = (char*)std::malloc((size_t)tot + 1);
	if (!o_buf) { nsh_set_error("nsh_map_batch: out of host memory"); return 0; }
	int64_t w = 0;
	for (const auto &s : text) std::memcpy(o_buf + w, s.data(), s.size()), w += (int64_t)s.size();
	o_buf[w] = 0;
	*out = o_buf, *out_len = w;
	*id += id0[n_q];
	for (int32_t q = 0; q < n_q; ++q) st[4] += n_out[q] > 0, st[5] += n_out[q];
	lap(5);
	if (stats) std::memcpy(stats, st, sizeof(st));
	return 1;
}
