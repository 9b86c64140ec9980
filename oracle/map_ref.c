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
