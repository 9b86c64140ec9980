/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: exposes the reference's file-static PSSM diagonal evidence
 * (mp_diag_seg and mp_chain_pssm_ev, map.c:91-139) by compiling its map.c where it lies
 * (REF is on the include path), so the library's batched kernel can be compared with it. */
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
