/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: runs the reference's own output path for one alignment so the
 * library's statistics and records can be compared with it byte for byte.
 *
 * The reference's align.c is compiled here, where it lies (REF is on the include path), for
 * its file-static mp_extra_gen / mp_extra_stop / mp_extra_start / mp_extra_cal
 * (align.c:148-306); mp_write_output comes from its format.c (format.c:457-477) and
 * mp_ntseq_get_by_v from its ntseq.c (ntseq.c:92-118). The genome is one contig packed the
 * way ntseq.c stores sequences (two 4-bit codes per byte, even positions in the low nibble). */
#include "align.c"

int64_t or_ref_output(const char *ctg_name, const uint8_t *ctg_nt4, int64_t ctg_len, const char *qname, const char *aa,
		int32_t qlen, int32_t rev, int32_t qs, int32_t qe, int64_t vs, int64_t ve, int64_t as, int64_t ae,
		const uint32_t *cigar, int32_t n_cigar, int32_t score, int32_t profile_score, int32_t go, int32_t ge, int32_t fs,
		const int8_t *mat, const int8_t *pssm_row, uint32_t flag, uint32_t acc_flags, int32_t rank_sc, int32_t chn_sc,
		const char *gff_prefix, int64_t id, int32_t gff_delim, int32_t hit_idx, char *out, int64_t size)
{
	mp_ntdb_t nt;
	mp_ctg_t ctg;
	mp_idx_t mi;
	mp_mapopt_t opt;
	mp_reg1_t r;
	mp_bseq1_t seq;
	mp_cigar_t cig;
	kstring_t s = {0, 0, 0};
	uint8_t *reg;
	int64_t i, l;

	memset(&nt, 0, sizeof(nt)), memset(&ctg, 0, sizeof(ctg)), memset(&mi, 0, sizeof(mi));
	memset(&opt, 0, sizeof(opt)), memset(&r, 0, sizeof(r)), memset(&seq, 0, sizeof(seq));
	nt.seq = (uint8_t*)calloc(ctg_len / 2 + 2, 1);
	for (i = 0; i < ctg_len; ++i) nt.seq[i >> 1] |= (ctg_nt4[i] & 0xf) << ((i & 1) * 4);
	nt.n_ctg = nt.m_ctg = 1, nt.l_seq = ctg_len;
	ctg.off = 0, ctg.len = ctg_len, ctg.name = (char*)ctg_name;
	nt.ctg = &ctg;
	mi.nt = &nt;
	opt.flag = flag, opt.acc_flags = acc_flags, opt.asize = 22, opt.go = go, opt.ge = ge, opt.fs = fs;
	opt.gff_prefix = gff_prefix, opt.gff_delim = gff_delim;
	memcpy(opt.mat, mat, 484);
	r.vid = rev & 1, r.qs = qs, r.qe = qe, r.vs = vs, r.ve = ve, r.rank_sc = rank_sc, r.chn_sc = chn_sc;
	reg = (uint8_t*)malloc(ae - as + 1);
	if (mp_ntseq_get_by_v(&nt, r.vid, as, ae, reg) != ae - as) { free(reg); free(nt.seq); return -1; }
	cig.n = cig.m = n_cigar;
	cig.c = (uint32_t*)malloc(n_cigar * 4 + 4);
	memcpy(cig.c, cigar, n_cigar * 4);
	mp_extra_gen(0, &r, &cig, score);          /* as align.c:625-629 */
	r.p->profile_score = profile_score;
	r.p->dist_stop  = mp_extra_stop(&r, reg, as, ae);
	r.p->dist_start = mp_extra_start(&r, reg, as, ae);
	mp_extra_cal(&r, &opt, &reg[vs - as], (ae - as) - (vs - as), &aa[qs], qlen, 0, pssm_row);
	seq.l_seq = qlen, seq.name = (char*)qname, seq.seq = (char*)aa;
	mp_write_output(&s, 0, &mi, &seq, &r, &opt, id, hit_idx);
	l = (int64_t)s.l;
	if (size > 0) {
		int64_t c = l < size - 1? l : size - 1;
		memcpy(out, s.s, c);
		out[c] = 0;
	}
	free(s.s), free(r.p), free(r.feat), free(reg), free(nt.seq);
	return l;
}

/* The reference's align stage (mp_align_batch, align.c:745-777, default options) on the regions of
 * one query as or_ref_align_inputs returns them, against an index from or_ref_index_build; then
 * each aligned region's record through mp_write_output (format.c:457-477; flag selects PAF or
 * GFF3). Per region: st[k] = 1 when aligned, 0 when dropped; fields[8k..] = vid, qs, qe,
 * vs, ve, dp_score, dp_max, n_cigar; the CIGARs go to cig (up to ccap ops in all), the records,
 * one after another, to text (up to size bytes). Returns the bytes of text written, -1 on overflow. */
typedef struct {
	int32_t vid, qs, qe, cnt, chn_sc, chn_sc_ungap, id, parent;
	int64_t vs, ve;
	uint64_t ext;
} or_areg2_t;
const mp_idx_t *or_ref_index_mi(void *handle);

int64_t or_ref_align_stage(void *idx_handle, const char *qname, const char *aa, int32_t qlen, int32_t n,
		const or_areg2_t *in, const uint64_t *anch, uint32_t flag, int32_t *st, int64_t *fields, int64_t ccap,
		uint32_t *cig, char *text, int64_t size)
{
	const mp_idx_t *mi = or_ref_index_mi(idx_handle);
	mp_mapopt_t opt;
	mp_reg1_t *reg, **rp;
	uint64_t *ext;
	mp_bseq1_t seq;
	kstring_t s = {0, 0, 0};
	int64_t k = 0, nc = 0, tl = 0, ret;
	int32_t i;
	mp_mapopt_init(&opt);
	opt.flag |= flag;
	reg = (mp_reg1_t*)calloc(n + 1, sizeof(mp_reg1_t));
	rp = (mp_reg1_t**)calloc(n + 1, sizeof(void*));
	ext = (uint64_t*)calloc(n + 1, 8);
	for (i = 0; i < n; ++i) {
		mp_reg1_t *r = &reg[i];
		r->vid = in[i].vid, r->qs = in[i].qs, r->qe = in[i].qe, r->cnt = in[i].cnt, r->chn_sc = in[i].chn_sc;
		r->chn_sc_ungap = in[i].chn_sc_ungap, r->id = in[i].id, r->parent = in[i].parent;
		r->vs = in[i].vs, r->ve = in[i].ve, r->off = 0;
		r->a = (uint64_t*)malloc((size_t)(r->cnt + 1) * 8);
		memcpy(r->a, &anch[k], (size_t)r->cnt * 8);
		k += r->cnt;
		ext[i] = in[i].ext, rp[i] = r;
	}
	mp_align_batch(0, &opt, mi, qlen, aa, n, rp, ext, 0, 0, st, 0);
	memset(&seq, 0, sizeof(seq));
	seq.l_seq = qlen, seq.name = (char*)qname, seq.seq = (char*)aa;
	for (i = 0; i < n; ++i) {
		const mp_reg1_t *r = &reg[i];
		int64_t *f = &fields[8 * i];
		st[i] = r->p != 0;   /* aligned, or dropped (mp_align_start found no trusted anchor) */
		f[0] = r->vid, f[1] = r->qs, f[2] = r->qe, f[3] = r->vs, f[4] = r->ve;
		f[5] = r->p? r->p->dp_score : 0, f[6] = r->p? r->p->dp_max : 0, f[7] = r->p? r->p->n_cigar : 0;
		if (!r->p) continue;
		if (nc + r->p->n_cigar <= ccap) memcpy(&cig[nc], r->p->cigar, (size_t)r->p->n_cigar * 4);
		nc += r->p->n_cigar;
		mp_write_output(&s, 0, mi, &seq, r, &opt, i, 0);   /* resets s: copy each record out */
		if (tl + (int64_t)s.l <= size) memcpy(text + tl, s.s, s.l);
		tl += (int64_t)s.l;
	}
	ret = nc > ccap || tl > size? -1 : tl;
	for (i = 0; i < n; ++i) { free(reg[i].p); free(reg[i].feat); free(reg[i].a); }
	free(reg), free(rp), free(ext), free(s.s);
	return ret;
}
