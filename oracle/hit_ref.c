/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: runs the reference's --gpu-prefilter candidate selection
 * (mp_gapless_calibrate, mp_gapless_adaptive_k and mp_select_gapless, hit.c:149-263) the way
 * mp_gpu_prefilter_batch and mp_gpu_prefilter_select call them (map.c:1075-1083, 507-521), by
 * compiling its hit.c where it lies (REF is on the include path). */
#include "hit.c"

/* n_query queries; query q owns candidates cand_off[q]..cand_off[q+1]-1 with gapless scores,
 * chain scores, vid and vs. keep[i] = 1 for the survivors; k_out[q] = the K handed to
 * mp_select_gapless (0 for a query without candidates); cal = the calibration (zero unless
 * adaptive). */
void or_ref_gapless_select(int32_t n_query, const int32_t *cand_off, const int32_t *qlen, const int32_t *scores,
		const int32_t *chn_sc, const uint32_t *vid, const int64_t *vs, int32_t top_k, int32_t min_per_vid,
		int32_t adaptive, uint8_t *keep, int32_t *k_out, mp_gapless_calib_t *cal)
{
	int32_t q, i;
	memset(cal, 0, sizeof(*cal));
	if (adaptive) mp_gapless_calibrate(n_query, cand_off, scores, cal);
	for (i = 0; i < cand_off[n_query]; ++i) keep[i] = 0;
	for (q = 0; q < n_query; ++q) {
		int32_t n = cand_off[q + 1] - cand_off[q], k = top_k;
		mp_reg1_t *reg;
		k_out[q] = 0;
		if (n == 0) continue;
		reg = (mp_reg1_t*)calloc(n, sizeof(mp_reg1_t));
		for (i = 0; i < n; ++i) {
			reg[i].id = cand_off[q] + i;
			reg[i].chn_sc = chn_sc[cand_off[q] + i], reg[i].vid = vid[cand_off[q] + i], reg[i].vs = vs[cand_off[q] + i];
		}
		if (adaptive) {
			if (k > MP_GAPLESS_ADAPT_MAX_K) k = MP_GAPLESS_ADAPT_MAX_K;
			k = mp_gapless_adaptive_k(n, &scores[cand_off[q]], qlen[q], k, cal);
		}
		k_out[q] = k;
		mp_select_gapless(0, k, min_per_vid, &n, reg, &scores[cand_off[q]]);
		for (i = 0; i < n; ++i) keep[reg[i].id] = 1;
		free(reg);
	}
}
