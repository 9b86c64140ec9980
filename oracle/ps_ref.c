/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: runs the reference's six-frame ORF walk (ps_orf_worker and
 * ps_add_orf, prescreen.c:87-132; file-static) by compiling its prescreen.c where it lies
 * (REF is on the include path), on a one-contig genome packed the way ntseq.c stores it. */
#include "prescreen.c"

/* ORFs of strand rev (0 forward, 1 reverse) of a contig of base codes 0..4: the count, and up to
 * max entries of (start, codons) and max_aa residue codes in the reference's emission order */
int64_t or_ref_orfs(const uint8_t *ctg_nt4, int64_t len, int32_t rev, int32_t min_orf, int64_t max, int64_t *st,
		int32_t *n_aa, int64_t max_aa, uint8_t *aa, int64_t *n_aa_total)
{
	mp_ntdb_t nt;
	mp_ctg_t ctg;
	ps_orf_v orf[2];
	ps_aa_v av[2];
	ps_orf_aux_t aux;
	int64_t i, n;
	memset(&nt, 0, sizeof(nt)), memset(&ctg, 0, sizeof(ctg)), memset(orf, 0, sizeof(orf)), memset(av, 0, sizeof(av));
	nt.seq = (uint8_t*)calloc(len / 2 + 2, 1);
	for (i = 0; i < len; ++i) nt.seq[i >> 1] |= (ctg_nt4[i] & 0xf) << ((i & 1) * 4);
	nt.n_ctg = nt.m_ctg = 1, nt.l_seq = len;
	ctg.off = 0, ctg.len = len, ctg.name = (char*)"c";
	nt.ctg = &ctg;
	aux.nt = &nt, aux.min_orf = min_orf, aux.orf = orf, aux.aa = av;
	ps_orf_worker(&aux, rev & 1, 0);
	n = orf[rev & 1].n;
	for (i = 0; i < n && i < max; ++i) st[i] = orf[rev & 1].a[i].vs, n_aa[i] = orf[rev & 1].a[i].aa_len;
	*n_aa_total = av[rev & 1].n;
	memcpy(aa, av[rev & 1].a, av[rev & 1].n < max_aa? av[rev & 1].n : max_aa);
	free(orf[rev & 1].a), free(av[rev & 1].a), free(nt.seq);
	return n;
}

/* ps_gapped (prescreen.c:346-382): banded local affine score of a profile against one ORF */
int32_t or_ref_ps_gapped(const int8_t *pssm, int32_t L, const uint8_t *orf, int32_t n, int32_t diag, int32_t band,
		int32_t go, int32_t ge)
{
	ps_scratch_t s;
	int32_t r;
	memset(&s, 0, sizeof(s));
	r = ps_gapped(&s, pssm, L, orf, n, diag, band, go, ge);
	free(s.H), free(s.F);
	return r;
}

/* The whole per-profile prescreen search of the reference: mp_prescreen_build over a genome of
 * contigs (base codes 0..4, forward strand, concatenated at ctg_off), then ps_query_worker
 * (prescreen.c:402-486) for one profile h (an mp_hmm_t from mp_hmm_read1). Writes up to max
 * candidates as mp_cand1_t (the worker's order, before mp_cand_tab_set merges them) and returns
 * how many there were; *n_orf / *n_aa receive the ORF database's size. */
int32_t or_ref_ps_search(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		const mp_hmm_t *h, int32_t kmer, int32_t min_orf, float sens, int32_t kmer_thr, int32_t kmer_max,
		int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge, double max_evalue,
		int32_t max, mp_cand1_t *out, int64_t *n_orf, int64_t *n_aa)
{
	mp_ntdb_t nt;
	mp_psopt_t po;
	mp_prescreen_t *ps;
	ps_batch_aux_t aux;
	int64_t i, tot = 0;
	int32_t c, n;
	memset(&nt, 0, sizeof(nt));
	for (c = 0; c < n_ctg; ++c) tot = ctg_off[c] + ctg_len[c] > tot? ctg_off[c] + ctg_len[c] : tot;
	nt.seq = (uint8_t*)calloc(tot / 2 + 2, 1);
	for (i = 0; i < tot; ++i) nt.seq[i >> 1] |= (nt4[i] & 0xf) << ((i & 1) * 4);
	nt.n_ctg = nt.m_ctg = n_ctg, nt.l_seq = tot;
	nt.ctg = (mp_ctg_t*)calloc(n_ctg, sizeof(mp_ctg_t));
	for (c = 0; c < n_ctg; ++c) nt.ctg[c].off = ctg_off[c], nt.ctg[c].len = ctg_len[c], nt.ctg[c].name = (char*)"c";
	mp_psopt_init(&po);
	po.kmer = kmer, po.min_orf = min_orf, po.sens = sens, po.kmer_thr = kmer_thr, po.kmer_max = kmer_max;
	po.min_ungapped = min_ungapped, po.max_seqs = max_seqs, po.band = band, po.go = go, po.ge = ge;
	ps = mp_prescreen_build(&nt, &po, 1);
	*n_orf = ps->n_orf, *n_aa = ps->n_aa;
	memset(&aux, 0, sizeof(aux));
	aux.ps = ps, aux.po = &po, aux.max_evalue = max_evalue, aux.thr = mp_ps_kmer_thr(&po);
	aux.hmm = (mp_hmm_t**)&h;
	aux.scr = (ps_scratch_t*)calloc(1, sizeof(ps_scratch_t));
	aux.res = (mp_cand1_t**)calloc(1, sizeof(mp_cand1_t*));
	aux.n_res = (int32_t*)calloc(1, sizeof(int32_t));
	ps_query_worker(&aux, 0, 0);
	n = aux.n_res[0];
	memcpy(out, aux.res[0], (size_t)(n < max? n : max) * sizeof(mp_cand1_t));
	free(aux.res[0]);
	ps_scratch_free(aux.scr);
	free(aux.scr), free(aux.res), free(aux.n_res);
	mp_prescreen_destroy(ps);
	free(nt.ctg), free(nt.seq);
	return n;
}

int32_t or_ref_ps_kmer_thr(int32_t kmer, float sens, int32_t kmer_thr)
{
	mp_psopt_t po;
	mp_psopt_init(&po);
	po.kmer = kmer, po.sens = sens, po.kmer_thr = kmer_thr;
	return mp_ps_kmer_thr(&po);
}

/* build once, search many (timing the reference's search on a large genome) */
typedef struct { mp_ntdb_t nt; mp_prescreen_t *ps; } or_ps_t;

void *or_ref_ps_build(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len, int32_t kmer,
		int32_t min_orf)
{
	or_ps_t *r = (or_ps_t*)calloc(1, sizeof(or_ps_t));
	mp_psopt_t po;
	int64_t i, tot = 0;
	int32_t c;
	for (c = 0; c < n_ctg; ++c) tot = ctg_off[c] + ctg_len[c] > tot? ctg_off[c] + ctg_len[c] : tot;
	r->nt.seq = (uint8_t*)calloc(tot / 2 + 2, 1);
	for (i = 0; i < tot; ++i) r->nt.seq[i >> 1] |= (nt4[i] & 0xf) << ((i & 1) * 4);
	r->nt.n_ctg = r->nt.m_ctg = n_ctg, r->nt.l_seq = tot;
	r->nt.ctg = (mp_ctg_t*)calloc(n_ctg, sizeof(mp_ctg_t));
	for (c = 0; c < n_ctg; ++c) r->nt.ctg[c].off = ctg_off[c], r->nt.ctg[c].len = ctg_len[c], r->nt.ctg[c].name = (char*)"c";
	mp_psopt_init(&po);
	po.kmer = kmer, po.min_orf = min_orf;
	r->ps = mp_prescreen_build(&r->nt, &po, 1);
	return r;
}

int32_t or_ref_ps_query(void *handle, const mp_hmm_t *h, int32_t kmer, float sens, int32_t kmer_max, int32_t min_ungapped,
		int32_t max_seqs, int32_t band, int32_t go, int32_t ge, double max_evalue, int32_t max, mp_cand1_t *out,
		int64_t *stats)
{
	or_ps_t *r = (or_ps_t*)handle;
	mp_psopt_t po;
	ps_batch_aux_t aux;
	int32_t n;
	mp_psopt_init(&po);
	po.kmer = kmer, po.sens = sens, po.kmer_max = kmer_max, po.min_ungapped = min_ungapped, po.max_seqs = max_seqs;
	po.band = band, po.go = go, po.ge = ge;
	memset(&aux, 0, sizeof(aux));
	aux.ps = r->ps, aux.po = &po, aux.max_evalue = max_evalue, aux.thr = mp_ps_kmer_thr(&po);
	aux.hmm = (mp_hmm_t**)&h;
	aux.scr = (ps_scratch_t*)calloc(1, sizeof(ps_scratch_t));
	aux.res = (mp_cand1_t**)calloc(1, sizeof(mp_cand1_t*));
	aux.n_res = (int32_t*)calloc(1, sizeof(int32_t));
	ps_query_worker(&aux, 0, 0);
	n = aux.n_res[0];
	memcpy(out, aux.res[0], (size_t)(n < max? n : max) * sizeof(mp_cand1_t));
	if (stats) {
		const ps_scratch_t *c = aux.scr;
		stats[0] = c->n_win, stats[1] = c->n_kmer, stats[2] = c->n_lookup, stats[3] = c->n_dbl;
		stats[4] = c->n_orf_scored, stats[5] = c->n_aligned, stats[6] = c->n_kept;
	}
	free(aux.res[0]);
	ps_scratch_free(aux.scr);
	free(aux.scr), free(aux.res), free(aux.n_res);
	return n;
}

void or_ref_ps_free(void *handle)
{
	or_ps_t *r = (or_ps_t*)handle;
	if (!r) return;
	mp_prescreen_destroy(r->ps);
	free(r->nt.ctg), free(r->nt.seq), free(r);
}
