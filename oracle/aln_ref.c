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
