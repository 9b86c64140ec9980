/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: the reference's k-mer block index (mp_idx_build without the file
 * read: mp_idx_boff, mp_sketch_nt4 per contig strand as build_worker calls it, build_bidx;
 * index.c:14-29, 52-97, sketch.c:108-169) on a genome given as arrays, by compiling its
 * index.c where it lies (REF is on the include path); and its chaining (mp_chain, chain.c,
 * compiled with the reference sources) on a copy of a query's anchors. */
#include "index.c"

typedef struct { mp_ntdb_t nt; mp_idx_t mi; } or_idx_t;

void *or_ref_index_build(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit)
{
	or_idx_t *r = (or_idx_t*)calloc(1, sizeof(or_idx_t));
	mp64_v *a = (mp64_v*)calloc(n_ctg * 2 + 1, sizeof(mp64_v));
	int64_t i, tot = 0;
	int32_t c, j;
	for (c = 0; c < n_ctg; ++c) tot = ctg_off[c] + ctg_len[c] > tot? ctg_off[c] + ctg_len[c] : tot;
	r->nt.seq = (uint8_t*)calloc(tot / 2 + 2, 1);
	for (i = 0; i < tot; ++i) r->nt.seq[i >> 1] |= (nt4[i] & 0xf) << ((i & 1) * 4);
	r->nt.n_ctg = r->nt.m_ctg = n_ctg, r->nt.l_seq = tot;
	r->nt.ctg = (mp_ctg_t*)calloc(n_ctg, sizeof(mp_ctg_t));
	for (c = 0; c < n_ctg; ++c) r->nt.ctg[c].off = ctg_off[c], r->nt.ctg[c].len = ctg_len[c], r->nt.ctg[c].name = (char*)"c";
	memset(&r->mi.opt, 0, sizeof(r->mi.opt));
	r->mi.opt.min_aa_len = min_aa_len, r->mi.opt.kmer = kmer, r->mi.opt.mod_bit = mod_bit, r->mi.opt.bbit = bbit;
	r->mi.nt = &r->nt;
	r->mi.bo = mp_idx_boff(&r->nt, bbit, &r->mi.n_block);
	for (j = 0; j < n_ctg * 2; ++j) {
		uint8_t *seq = (uint8_t*)malloc(ctg_len[j >> 1] + 1);
		int64_t len = mp_ntseq_get(&r->nt, j >> 1, 0, -1, j & 1, seq);
		mp64_v s = {0, 0, 0};
		mp_sketch_nt4(0, seq, len, min_aa_len, kmer, mod_bit, bbit, r->mi.bo[j], &s);
		a[j] = s;
		free(seq);
	}
	build_bidx(&r->mi.opt, &r->mi, a);
	for (j = 0; j < n_ctg * 2; ++j) free(a[j].a);
	free(a);
	return r;
}

/* n_kb; copies the bucket starts ki (n_bucket), the entries kb (n_kb) and the block offsets bo
 * (2 n_ctg + 1) into the arrays that are not NULL */
int64_t or_ref_index_get(void *handle, int64_t *ki, uint32_t *kb, uint32_t *bo)
{
	or_idx_t *r = (or_idx_t*)handle;
	if (ki) memcpy(ki, r->mi.ki, (size_t)mp_n_bucket(&r->mi.opt) * sizeof(int64_t));
	if (kb) memcpy(kb, r->mi.kb, (size_t)r->mi.n_kb * sizeof(uint32_t));
	if (bo) memcpy(bo, r->mi.bo, (size_t)(r->nt.n_ctg * 2 + 1) * sizeof(uint32_t));
	return r->mi.n_kb;
}

void or_ref_index_free(void *handle)
{
	or_idx_t *r = (or_idx_t*)handle;
	if (!r) return;
	free(r->mi.ki), free(r->mi.kb), free(r->mi.bo);
	free(r->nt.ctg), free(r->nt.seq), free(r);
}

/* the index inside a handle, for the anchor adaptor in map_ref.c */
const mp_idx_t *or_ref_index_mi(void *handle)
{
	return &((or_idx_t*)handle)->mi;
}

/* mp_chain (chain.c:160-219) on a copy of n anchors: returns the number of chains and writes
 * their u[] (score << 32 | anchors, up to max_u) and their anchors (up to max_a), in the order
 * mp_chain returns them */
int32_t or_ref_chain(int32_t max_dist_x, int32_t max_dist_y, int32_t bw, int32_t max_skip, int32_t max_iter,
		int32_t min_cnt, int32_t min_sc, float chn_coef_log, int32_t is_spliced, int32_t kmer, int32_t bbit, int64_t n,
		const uint64_t *a, int32_t max_u, uint64_t *u_out, int64_t max_a, uint64_t *a_out)
{
	uint64_t *b = (uint64_t*)malloc((size_t)(n > 0? n : 1) * 8), *u = 0, *r;
	int32_t n_u = 0, i;
	int64_t k = 0;
	if (n > 0) memcpy(b, a, (size_t)n * 8);
	r = mp_chain(max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc, chn_coef_log, is_spliced, kmer, bbit,
			n, n > 0? b : 0, &n_u, &u, 0);
	if (n == 0) free(b);
	for (i = 0; i < n_u; ++i) {
		if (i < max_u) u_out[i] = u[i];
		k += (uint32_t)u[i];
	}
	for (i = 0; i < k && i < max_a; ++i) a_out[i] = r[i];
	free(u), free(r);
	return n_u;
}
