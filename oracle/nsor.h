/*
 * nsor.h -- the oracle: a plain-C restatement of the reference's CPU semantics for the
 * spliced-DP hot path. TEST INFRASTRUCTURE ONLY: it is linked by tests/, bench.py's CPU
 * baseline leg and __graft_entry__.smoke() as the checker, never by the product library.
 *
 * Every function cites the reference file:line it restates. Parity pinning (what checks
 * each function against the compiled reference, golden vectors or closed forms) is listed
 * in oracle/README.md.
 */
#ifndef NSOR_H
#define NSOR_H
#include <stdint.h>
#include "../include/nasw-hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t nt_len, aa_len, score; } or_ext_t;

/* tables (nasw-tab.c:6-107) */
extern uint8_t or_nt4[256], or_aa20[256], or_codon[64];
extern int8_t or_blosum62[484];
int or_make_tables(int codon_type);
void or_opt_init(ns_opt_t *opt);                 /* nasw-tab.c:131-147 */
void or_opt_set_sp(ns_opt_t *opt, int32_t model); /* nasw-tab.c:124-129 */
void or_set_stop_sc(int32_t asize, int8_t *mat, int8_t pen); /* nasw-tab.c:149-156 */
int32_t or_ie_penalty(float ie_coef, int32_t d); /* nasw-sse.c:341-349, 758-761 */

/* input preparation (nasw-sse.c:95-214); buffers: nas[nl], aas[al], donor[nl+1], acceptor[nl+1] */
void or_prep(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int left, uint8_t *nas, uint8_t *aas, int8_t *donor, int8_t *acceptor);

/* 16-bit score-only DP over npad columns (nasw-sse.c:360-459, 596-741): extension mode when
 * opt->flag has EXT_LEFT/RIGHT, plain score-only global otherwise. rows_done (may be NULL)
 * receives the number of DP rows evaluated (last row index - 1). */
void or_dp16_score(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t npad, or_ext_t *out, int32_t *rows_done);

/* ns_global_gs16b / ns_global_gs32b including the CIGAR path (nasw-sse.c:351-567, 779-973,
 * 44-93): a lane-by-lane emulation of the striped SIMD evaluation with vsize lanes
 * (8 for int16, 4 for int32), reproducing the trace bits and the backtrack exactly.
 * r->cigar is malloc'ed (free with free()). npad_ext: stripe width used in extension mode. */
void or_global16(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t ext_width, ns_rst_t *r);
void or_global32(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, ns_rst_t *r);
void or_global_auto(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t ext_width, ns_rst_t *r); /* nasw-sse.c:980-999 (CPU branch) */
/* raw trace words of the CIGAR path, [nl][slen*vsize] in column order (tests of the GPU trace) */
int32_t or_global_trace(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t is32, uint16_t *trace_out, int32_t *score_out);

/* HMM profile (hmm.c:46-81, 112-195, 252-291) */
typedef struct {
	int32_t len;
	int16_t *emit;  /* [len][20], HMMER alphabet order, best residue at +4 */
	int16_t *trans; /* [len][7] m->m m->i m->d i->m i->i d->m d->d */
	int8_t *pssm;   /* [len][22] */
} or_hmm_t;
int or_hmm_parse(const char *text, or_hmm_t *h, const char **next); /* one HMMER3 record */
void or_hmm_free(or_hmm_t *h);
int32_t or_hmm_viterbi(const or_hmm_t *h, const uint8_t *nt, int32_t nlen);
const int8_t *or_hmm_aa20_to_alpha(void);

/* gapless X-drop candidate score (nasw-cuda.cu:1151-1295); one candidate */
int32_t or_gapless_score(const uint8_t *packed_nt, const uint8_t *queries,
		const ns_gpu_gapless_candidate_t *c, const ns_gpu_gapless_anchor_t *anchors,
		int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *mat, int32_t asize, const uint8_t *codon);

/* PSSM ungapped diagonal segment, Kadane with clamp at 0 (prescreen.c:334-344) */
int32_t or_pssm_ungapped(const int8_t *pssm, int32_t plen, const uint8_t *orf, int32_t olen,
		int32_t diag, int32_t *qs, int32_t *qe);

#ifdef __cplusplus
}
#endif
#endif
