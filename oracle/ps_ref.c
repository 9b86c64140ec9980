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
