/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: exposes the reference's file-static PSSM diagonal evidence
 * (mp_diag_seg and mp_chain_pssm_ev, map.c:91-139) and anchor collection (map.c:583-597) by
 * compiling its map.c where it lies (REF is on the include path), so the library's batched
 * kernels can be compared with them. */
#include "map.c"

int32_t or_ref_diag_seg(const uint8_t *nt, int64_t l_nt, const int8_t *pssm, int32_t qlen, int32_t kmer, int64_t x,
		int32_t q, int32_t xdrop, int32_t *aa_e)
{
	return mp_diag_seg(nt, l_nt, pssm, qlen, kmer, x, q, xdrop, aa_e);
}

int32_t or_ref_chain_pssm_ev(const uint8_t *nt, int64_t l_nt, const int8_t *pssm, int32_t qlen, int32_t kmer, int32_t n,
		const uint64_t *a, int32_t min_seg, int32_t xdrop)
{
	return mp_chain_pssm_ev(nt, l_nt, pssm, qlen, kmer, n, a, min_seg, xdrop);
}

/* the anchors of one query (ASCII residues) against an index from or_ref_index_build, as
 * mp_map_coarse collects them with the default options (map.c:583-597: mp_sketch_query,
 * mp_cal_max_occ capped at max_occ, mp_collect_anchors, radix_sort_mp64); returns their count
 * and writes up to max of them */
const mp_idx_t *or_ref_index_mi(void *handle);
int64_t or_ref_anchors(void *idx_handle, const char *seq, int32_t qlen, int32_t max_occ, int64_t max, uint64_t *out)
{
	const mp_idx_t *mi = or_ref_index_mi(idx_handle);
	mp_mapopt_t opt;
	mp64_v sd = {0, 0, 0};
	int64_t n_a = 0, n_rescued = 0, i;
	int32_t occ;
	uint64_t *a;
	memset(&opt, 0, sizeof(opt));
	opt.max_occ = max_occ;
	mp_sketch_query(0, &opt, seq, qlen, mi->opt.kmer, mi->opt.mod_bit, 0, &sd);
	if (sd.n >= 8) {
		occ = mp_cal_max_occ(0, mi, sd.n, sd.a);
		if (occ > opt.max_occ) occ = opt.max_occ;
	} else occ = opt.max_occ;
	a = mp_collect_anchors(0, mi, &opt, &sd, occ, "q", &n_a, &n_rescued);
	free(sd.a);
	radix_sort_mp64(a, a + n_a);
	for (i = 0; i < n_a && i < max; ++i) out[i] = a[i];
	free(a);
	return n_a;
}

/* The inputs of the align stage for one query, as the reference's default map path builds them
 * (mp_mapopt_init defaults, as main() leaves them without -I/-G): mp_map_coarse (map.c:547-641),
 * then the refine half of mp_map_finish (map.c:655-700: query kmer2 sketch, mp_cal_max_ext,
 * mp_refine_reg, mp_collate_a, mp_sort_reg, mp_set_parent, mp_select_sub) and the extension limits
 * mp_map_finish hands to mp_align_batch (map.c:709-716). Writes up to cap regions (their anchors
 * one after another into anch, up to acap) and returns the region count (-1 when a cap is exceeded). */
typedef struct {
	int32_t vid, qs, qe, cnt, chn_sc, chn_sc_ungap, id, parent;
	int64_t vs, ve;
	uint64_t ext;      /* extl << 32 | extr */
} or_areg_t;

int32_t or_ref_align_inputs(void *idx_handle, const char *qname, const char *seq, int32_t qlen, int32_t cap,
		or_areg_t *out, int64_t acap, uint64_t *anch, int64_t *n_anch)
{
	const mp_idx_t *mi = or_ref_index_mi(idx_handle);
	mp_mapopt_t opt;
	mp_tbuf_t *b;
	mp_coarse_map_t c;
	mp_reg1_t *reg;
	uint64_t *a, *ext;
	void *km, *a_km;
	int32_t i, n_reg, nr = 0, ret;
	int64_t k = 0;
	mp64_v sd_aa = {0, 0, 0};
	mp_mapopt_init(&opt);
	b = mp_tbuf_init();
	km = b->km;
	mp_map_coarse(mi, qlen, seq, b, &opt, qname, &c);
	reg = c.reg, a = c.a, a_km = c.a_km, n_reg = c.n_reg;
	mp_sketch_query(km, &opt, seq, qlen, opt.kmer2, 0, opt.hmm_seed_max2, &sd_aa);
	ext = mp_cal_max_ext(km, 0, n_reg, reg, a, 100, opt.max_ext, opt.max_ext);
	for (i = 0; i < n_reg; ++i) {
		int32_t extl = ext? ext[i]>>32 : opt.max_ext, extr = ext? (int32_t)ext[i] : opt.max_ext;
		mp_refine_reg(km, mi, &opt, &sd_aa, &reg[i], extl, extr, 0, qlen, 0);
		if (reg[i].cnt > 0) reg[nr++] = reg[i];
	}
	n_reg = nr;
	kfree(km, sd_aa.a);
	kfree(km, ext);
	kfree(a_km, a);
	a = mp_collate_a(km, n_reg, reg);
	mp_sort_reg(km, &n_reg, reg, 0);
	mp_set_parent(km, opt.mask_level, opt.mask_len, n_reg, reg, mi->opt.kmer, 0);
	mp_select_sub(km, opt.pri_ratio * opt.pri_ratio, mi->opt.kmer * 2, c.select_n, &n_reg, reg);
	ext = mp_cal_max_ext(km, mi->nt, n_reg, reg, a, 100, opt.max_intron / 2, opt.max_intron / 2);
	ret = n_reg <= cap? n_reg : -1;
	for (i = 0; i < n_reg && ret >= 0; ++i) {
		const mp_reg1_t *r = &reg[i];
		or_areg_t *o = &out[i];
		o->vid = r->vid, o->qs = r->qs, o->qe = r->qe, o->cnt = r->cnt, o->chn_sc = r->chn_sc;
		o->chn_sc_ungap = r->chn_sc_ungap, o->id = r->id, o->parent = r->parent, o->vs = r->vs, o->ve = r->ve;
		o->ext = ext? ext[i] : 0;
		if (k + r->cnt > acap) { ret = -1; break; }
		memcpy(&anch[k], r->a, (size_t)r->cnt * 8);
		k += r->cnt;
	}
	*n_anch = k;
	for (i = 0; i < n_reg; ++i) { free(reg[i].p); free(reg[i].feat); }
	kfree(km, ext);
	kfree(km, a);
	free(reg);   /* mp_reg_gen_from_block allocates the regions from the heap (km = 0) */
	mp_tbuf_destroy(b);
	return ret;
}
