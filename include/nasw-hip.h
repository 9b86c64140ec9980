/*
 * nasw-hip.h -- C-ABI of the MI355X spliced-DP core.
 *
 * Drop-in boundary for the reference's DP layer: every type and constant below has
 * the same layout and value as the reference header (nasw.h:33-82, nasw.h:157-184),
 * and every ns_* entry point has the same signature, argument meaning and error
 * behaviour as the reference function it replaces (cited per declaration). A caller
 * written against nasw.h (align.c, map.c, hmm.c) links against libnasw_hip.so instead
 * of nasw-sse.o/nasw-tab.o/nasw-opencl.o/nasw-ext-gpu.o/hmm-gpu.o without source changes.
 *
 * All DP work runs on the GPU (hand-written gfx950 kernels); there is no CPU DP
 * path in this library. Entry points that the reference lets "fail back to the CPU"
 * (return 0 / NULL / INT32_MIN) only do so when no GPU is usable, and then report why
 * through ns_gpu_error().
 *
 * The nsh_* entry points at the end have no counterpart in the reference: they expose
 * the device-resident batch engine (inputs already in HBM) used by the benchmarks.
 */
#ifndef NASW_HIP_H
#define NASW_HIP_H

#include <stdint.h>
#include <stddef.h>
#include <string.h>

/* CIGAR operators (nasw.h:33-42) */
#define NS_CIGAR_M	0
#define NS_CIGAR_I	1
#define NS_CIGAR_D	2
#define NS_CIGAR_N	3  /* phase-0 intron */
#define NS_CIGAR_F	10 /* frameshift deletion */
#define NS_CIGAR_G	11 /* frameshift match */
#define NS_CIGAR_U	12 /* phase-1 intron */
#define NS_CIGAR_V	13 /* phase-2 intron */
#define NS_CIGAR_E	14 /* placeholder, unused */
#define NS_CIGAR_STR   "MIDNSHP=XBFGUVE"

/* option flags and splice models (nasw.h:44-53) */
#define NS_F_CIGAR      0x1
#define NS_F_EXT_LEFT   0x2
#define NS_F_EXT_RIGHT  0x4
#define NS_F_GPU        0x8

#define NS_S_NONE       0
#define NS_S_GENERIC    1
#define NS_S_MAMMAL     2

#define NS_SPSC_OFFSET 64

#ifdef __cplusplus
extern "C" {
#endif

/* lookup tables (nasw.h:55-58; filled by ns_make_tables) */
extern char *ns_tab_nt_i2c, *ns_tab_aa_i2c;
extern uint8_t ns_tab_a2r[22], ns_tab_nt4[256], ns_tab_aa20[256], ns_tab_aa13[256];
extern uint8_t ns_tab_codon[64], ns_tab_codon13[64];
extern int8_t ns_mat_blosum62[484];

/* alignment options; field order and types as nasw.h:62-75 */
typedef struct {
	int32_t flag;
	int32_t go, ge, io, fs;     /* gap open, gap extension, intron open, frameshift */
	int32_t xdrop, end_bonus;   /* extension x-drop; bonus for reaching the protein end */
	int32_t asize;              /* alphabet size (22) */
	int32_t gpu_device;
	int64_t gpu_min_cells;
	int32_t sp[6];              /* splice penalties: 0 pos3, 1 GC-AC, 2 AT-AC, 3 other, 4 pos0, 5 poly-Y */
	int32_t sp_null_bonus;      /* negative */
	float ie_coef;
	const int8_t *sc;           /* asize x asize substitution matrix */
	const int8_t *pssm;         /* optional al x asize position-specific scores */
	uint8_t *nt4, *aa20, *codon;
} ns_opt_t;

/* result; layout as nasw.h:77-82 */
typedef struct {
	int32_t n_cigar, m_cigar;
	int32_t nt_len, aa_len;
	int32_t score;
	uint32_t *cigar;            /* grown with krealloc(km, ...), freed by the caller with kfree(km, ...) */
} ns_rst_t;

/* ---- tables and options (nasw-tab.c) ---- */
int ns_make_tables(int codon_type);                                   /* nasw-tab.c:85-107 */
void ns_opt_init(ns_opt_t *opt);                                      /* nasw-tab.c:131-147 */
void ns_opt_set_sp(ns_opt_t *opt, int32_t model);                     /* nasw-tab.c:124-129 */
void ns_set_stop_sc(int32_t asize, int8_t *mat, int8_t score);        /* nasw-tab.c:149-156 */

/* ---- single-problem DP (nasw-sse.c); each call runs one GPU launch ---- */
void ns_global_gs16(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, ns_rst_t *r);  /* nasw-sse.c:774-777 */
void ns_global_gs32(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, ns_rst_t *r);  /* nasw-sse.c:975-978 */
void ns_global_gs16b(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r); /* nasw-sse.c:763-772 */
void ns_global_gs32b(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r); /* nasw-sse.c:779-973 */
void ns_global_gs_auto(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r); /* nasw-sse.c:980-999 */

/* ---- GPU global-DP API (nasw.h:140-152; nasw-cuda.cu:516-611, 1068-1145) ---- */
typedef struct ns_gpu_request_s ns_gpu_request_t;
int ns_gpu_init(int32_t device);
const char *ns_gpu_error(void);
const char *ns_gpu_name(void);
uint64_t ns_gpu_dispatch_count(void);
uint64_t ns_gpu_launch_count(void);
double ns_gpu_h2d_seconds(void);
double ns_gpu_kernel_seconds(void);
double ns_gpu_d2h_seconds(void);
ns_gpu_request_t *ns_gpu_submit(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss);
int ns_gpu_wait(ns_gpu_request_t *request, ns_rst_t *r);
int ns_gpu_global(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r);

/* ---- gapless X-drop prefilter (nasw.h:157-184; nasw-cuda.cu:1334-1412) ---- */
typedef struct {
	int64_t target_offset, target_length;
	int64_t query_offset;
	int32_t query_length;
	int32_t anchor_offset, anchor_count;
	int32_t reverse;
} ns_gpu_gapless_candidate_t;
typedef struct {
	int64_t target_block;
	int32_t query_end;
	int32_t reserved;
} ns_gpu_gapless_anchor_t;
typedef struct {
	uint64_t batches, candidates, anchors, seed_positions;
} ns_gpu_gapless_profile_t;
void ns_gpu_gapless_profile(ns_gpu_gapless_profile_t *profile);
int ns_gpu_gapless_supported(void);
void ns_gpu_gapless_reset_target(void);
int ns_gpu_gapless_score_batch(const uint8_t *packed_nt, int64_t packed_nt_bases,
		const uint8_t *queries, int64_t query_aa,
		const ns_gpu_gapless_candidate_t *candidates, int32_t n_candidates,
		const ns_gpu_gapless_anchor_t *anchors, int64_t n_anchors,
		int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *score_matrix, int32_t asize, const uint8_t *codon,
		int32_t *scores);

/* ---- --gpu-prefilter candidate selection on the device ----
 * nsh_gapless_calib_t has mp_gapless_calib_t's layout (mppriv.h:97-101).
 * nsh_gapless_select_batch: what mp_gpu_prefilter_batch does with the scores of one mini-batch
 * (map.c:1075-1083): with adaptive != 0, mp_gapless_calibrate over the batch (hit.c:149-189) and,
 * per query, K = mp_gapless_adaptive_k(n, scores, qlen, min(top_k, 4), cal) (hit.c:195-215);
 * otherwise K = top_k. Then mp_select_gapless(K, min_per_vid) (hit.c:217-263) per query, as
 * mp_gpu_prefilter_select runs it (map.c:507-521). Query q owns candidates
 * cand_off[q]..cand_off[q+1]-1 (at most 8192), with their gapless scores, chain scores (chn_sc),
 * vid and vs. keep[i] = 1 for the survivors. The caller compacts its regions in order, sets
 * chn_sc_ungap = scores[i] on all of them, then runs mp_sort_reg and mp_set_parent as map.c:523-524
 * does. k_out[q] (optional) receives K, 0 for a query without candidates; cal (optional) receives
 * the calibration, zero unless adaptive. Requires top_k >= 1 and 0 <= min_per_vid <= top_k
 * (options.c:174-181). Returns 1 on success, 0 on error.
 * nsh_gapless_prefilter_batch: ns_gpu_gapless_score_batch and the selection in one call. The
 * scores stay on the device between the two; they are also returned in scores. */
typedef struct { int32_t noise, n_query; float dup_frac; } nsh_gapless_calib_t;
int nsh_gapless_select_batch(int32_t n_query, const int32_t *cand_off, const int32_t *qlen, const int32_t *scores,
		const int32_t *chn_sc, const uint32_t *vid, const int64_t *vs, int32_t top_k, int32_t min_per_vid,
		int32_t adaptive, uint8_t *keep, int32_t *k_out, nsh_gapless_calib_t *cal);
int nsh_gapless_prefilter_batch(const uint8_t *packed_nt, int64_t packed_nt_bases,
		const uint8_t *queries, int64_t query_aa,
		const ns_gpu_gapless_candidate_t *candidates, int32_t n_candidates,
		const ns_gpu_gapless_anchor_t *anchors, int64_t n_anchors,
		int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *score_matrix, int32_t asize, const uint8_t *codon,
		int32_t n_query, const int32_t *cand_off, const int32_t *qlen, const int32_t *chn_sc, const uint32_t *vid,
		const int64_t *vs, int32_t top_k, int32_t min_per_vid, int32_t adaptive,
		int32_t *scores, uint8_t *keep, int32_t *k_out, nsh_gapless_calib_t *cal);

/* ---- profile-HMM Viterbi on the GPU (mppriv.h:128-134; hmm-gpu.c:282-352) ----
 * The profile object belongs to the caller's hmm.c (accessors mp_hmm_length, mp_hmm_emissions,
 * mp_hmm_transitions, mp_hmm_gpu_slot); its device copy is cached in gpu slot 0.
 * mp_hmm_dp_score_gpu_batch: score[i] = mp_hmm_dp_score(hmm, nt + off[i], len[i]); 1 on success.
 * mp_hmm_dp_score_gpu: INT32_MIN when nlen < 3 or the GPU path failed. */
struct mp_hmm_s;
int mp_hmm_dp_score_gpu_batch(const struct mp_hmm_s *hmm, int32_t n, const uint8_t *nt, int64_t l_nt, const int32_t *off, const int32_t *len, int32_t *score);
int32_t mp_hmm_dp_score_gpu(const struct mp_hmm_s *hmm, const uint8_t *nt, int32_t nlen);
void mp_hmm_gpu_release(struct mp_hmm_s *hmm);

/* ---- extension building blocks (nasw-sse.c:744-761) ---- */
uint8_t *ns_ext_prep(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, uint8_t **aas, int8_t **donor, int8_t **acceptor);
int32_t ns_ext_stripe_width(const ns_opt_t *opt, int32_t al);
int32_t ns_ext_ie_penalty(const ns_opt_t *opt, int32_t d);

/* ---- batched extension DP (nasw.h:196-212; nasw-ext-gpu.c:931-1114) ---- */
typedef struct ns_ext_batch_s ns_ext_batch_t;
ns_ext_batch_t *ns_ext_batch_init(void *km);
int32_t ns_ext_batch_add(ns_ext_batch_t *b, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss);
int32_t ns_ext_batch_flush(ns_ext_batch_t *b);
int ns_ext_batch_ready(ns_ext_batch_t *b, int32_t id);
void ns_ext_batch_wait(ns_ext_batch_t *b);
const ns_rst_t *ns_ext_batch_get(ns_ext_batch_t *b, int32_t id);
void ns_ext_batch_destroy(ns_ext_batch_t *b);
uint64_t ns_ext_gpu_launch_count(void);
uint64_t ns_ext_gpu_request_count(void);

/* kalloc ABI of the caller (kalloc.h:14-32): km == NULL means libc malloc. The library carries
 * weak fall-backs for km == NULL; a caller linking the reference's kalloc.c overrides them. */
void *kmalloc(void *km, size_t size);
void *krealloc(void *km, void *ptr, size_t size);
void *kcalloc(void *km, size_t count, size_t size);
void kfree(void *km, void *ptr);

/* CIGAR append with run merging (F and G never merge); same semantics as nasw.h:219-229 */
static inline uint32_t *ns_push_cigar(void *km, int32_t *n_cigar, int32_t *m_cigar, uint32_t *cigar, uint32_t op, int32_t len)
{
	if (*n_cigar == 0 || op != (cigar[(*n_cigar) - 1]&0xf) || op == NS_CIGAR_F || op == NS_CIGAR_G) {
		if (*n_cigar == *m_cigar) {
			(*m_cigar) += ((*m_cigar)>>1) + 8;
			cigar = (uint32_t*)krealloc(km, cigar, (size_t)(*m_cigar) * sizeof(uint32_t));
		}
		cigar[(*n_cigar)++] = (uint32_t)len<<4 | op;
	} else cigar[(*n_cigar)-1] += (uint32_t)len<<4;
	return cigar;
}
static inline void ns_rst_init(ns_rst_t *r) { memset(r, 0, sizeof(*r)); }

/* ======================================================================================
 * Device-resident batch engine (no counterpart in the reference).
 *
 * A plan owns HBM copies of many problems' raw inputs, sorted into length buckets.
 * nsh_ext_plan_run() prepares (codon translation, splice signals, query profiles) and
 * solves every problem with one persistent launch per bucket group; only the 12-byte
 * results travel back. Results equal ns_global_gs16b(flag = EXT_LEFT|EXT_RIGHT) of the
 * same request exactly.
 * ====================================================================================== */
typedef struct nsh_ext_plan_s nsh_ext_plan_t;
typedef struct { int32_t nt_len, aa_len, score; } nsh_ext_result_t;

/* All problems of a plan share opt (except the EXT_LEFT/EXT_RIGHT bit, taken per problem
 * from dirs[i] when dirs != NULL). With dirs == NULL and no EXT bit in opt->flag the plan runs
 * the plain score-only global DP instead (results equal ns_global_gs16b without CIGAR). ns/ss are concatenated raw nt strings (ASCII or 0..4),
 * as the concatenated proteins; ss may be NULL. Returns NULL and sets ns_gpu_error() on failure. */
nsh_ext_plan_t *nsh_ext_plan_create(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl,
		const char *as, const int64_t *as_off, const int32_t *al, const uint8_t *ss,
		const uint8_t *dirs, const ns_opt_t *opt);
int nsh_ext_plan_run(nsh_ext_plan_t *p);                      /* enqueue on the plan's stream; 1 ok, 0 error */
/* the same run as one hipGraph: captured on the first call (and again after the autotuner changed a
 * persistent grid), replayed by later calls; same results. No reference counterpart: for callers
 * that replay one resident batch many times. nsh_ext_plan_last_ms then covers the whole graph. */
int nsh_ext_plan_run_graph(nsh_ext_plan_t *p);
int nsh_ext_plan_sync(nsh_ext_plan_t *p);                     /* wait for the last run */
/* run only the problems with keep[i] != 0 (input order); the others report nt_len = aa_len = 0,
 * score INT32_MIN and no evaluated rows */
int nsh_ext_plan_run_subset(nsh_ext_plan_t *p, const uint8_t *keep);
/* the 64 prescreen features of every problem (input order), computed on the device from the
 * plan's resident inputs (also copied to host_out, n x 64, when it is not NULL); returns a device
 * pointer to n x 64 floats owned by the plan, or NULL (protein > 4096 aa or DNA > 24000 nt) */
const float *nsh_ext_plan_features(nsh_ext_plan_t *p, float *host_out);
int nsh_ext_plan_results(nsh_ext_plan_t *p, nsh_ext_result_t *out); /* copy results (input order) */
double nsh_ext_plan_cells(const nsh_ext_plan_t *p, int32_t kind);   /* 0: sum nl*al, 1: sum (rows done)*al after a run, 2: sum (rows done) */
float nsh_ext_plan_last_ms(nsh_ext_plan_t *p, int32_t which);  /* 0: whole run, 1: DP kernels only (HIP events) */
/* Autotuner: picks each column-count group's persistent wave count (waves per CU, one of the
 * compiled default's 1/2, 3/4, 1, 3/2 and 2 multiples), then the wavefronts per problem: whether
 * problems of at most 256 padded columns share a wave two at a time for every length, only up to
 * 16,384 or 65,536 rows, or never. Each choice is made by timing full runs of this plan, one at a
 * time (reps runs per candidate, the fastest of them; another setting replaces the default only
 * when it is 2% faster). Results do not depend on it.
 * waves_per_cu (may be NULL) receives the chosen values, one per group (3). 1 ok, 0 error. */
int nsh_ext_plan_autotune(nsh_ext_plan_t *p, int32_t reps, int32_t *waves_per_cu);
/* the row limit up to which small problems share a wave (INT32_MAX: all of them, 0: none) */
int32_t nsh_ext_plan_pair_rows(const nsh_ext_plan_t *p);
void nsh_ext_plan_destroy(nsh_ext_plan_t *p);

/* The align stage of the map path for many candidate regions at once (no single reference
 * counterpart: the reference runs it per query through mp_align_batch, align.c:745-777, with
 * mp_mapopt_t; this is its default-option path). Each candidate is a refined region as mp_map_finish
 * hands it over: its anchors (refine layout: target end of the k-mer relative to vs << 32 | query end)
 * and mp_cal_max_ext's limits. Every stage of mp_align_step (trusted-anchor filter, left extension and
 * io_end retry, gap DPs, right extension and retry, DP over the right extension) runs for all
 * candidates together on the device engines. */
typedef struct {
	int32_t go, ge, io, io_end, fs, xdrop, end_bonus;
	int32_t max_ext, max_intron, kmer2, end_retry_win;   /* mp_mapopt_t max_ext, max_intron, kmer2, acc_end_retry_win */
	float ie_coef, sp_scale;
	int32_t sp_model, sp_null_bonus;
	int8_t mat[484];                                    /* mp_mapopt_t.mat: BLOSUM62 with the stop scores set */
} nsh_align_opt_t;
typedef struct {
	int32_t query;          /* index into the query arrays */
	uint32_t vid;           /* contig << 1 | strand */
	int32_t qs, qe;         /* in: the refined chain's span; out: the alignment's */
	int64_t vs, ve;         /* same, in the strand's coordinates */
	int32_t cnt;            /* anchors anchors[a_off .. a_off + cnt) ; set to 0 when the candidate is dropped */
	int32_t chn_sc, chn_sc_ungap;
	int64_t a_off;
	uint64_t ext;           /* mp_cal_max_ext: extl << 32 | extr, 0 for none */
	/* out */
	int32_t status;         /* 1 aligned, 0 dropped (no trusted anchor) */
	int32_t dp_score, n_cigar;
	int64_t cigar_off;      /* into the CIGAR array */
	int64_t win_as, win_ae; /* the search window the statistics read */
} nsh_align_reg_t;
void nsh_align_opt_default(nsh_align_opt_t *o);    /* mp_mapopt_init's values (options.c:42-138) */
/* nt4: the genome as base codes 0..4 (contig c at ctg_off[c], ctg_len[c] bases); aa/qlen: the
 * queries (ASCII). *cigar receives one malloc'd array with every aligned candidate's CIGAR. 1 ok. */
int nsh_align_stage(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		const char *const *aa, const int32_t *qlen, int32_t n_reg, nsh_align_reg_t *reg, const uint64_t *anchors,
		const nsh_align_opt_t *opt, uint32_t **cigar);
/* The records of one aligned candidate (mp_extra_* and mp_write_output as mp_align_finish and the map
 * output write them; flag as nsh_write_output, id and hit_idx as the caller numbers its hits).
 * Returns the length (writes at most size - 1 bytes and a NUL), 0 for a dropped candidate, -1 on error. */
int64_t nsh_align_records(int32_t n_ctg, const uint8_t *nt4, const int64_t *ctg_off, const int64_t *ctg_len,
		const char *const *ctg_name, const char *const *qname, const char *const *aa, const int32_t *qlen,
		const nsh_align_reg_t *r, const uint32_t *cigar, const nsh_align_opt_t *opt, uint32_t flag, int64_t id,
		int32_t hit_idx, char *out, int64_t size);

/* The map path for a batch of proteins (what the reference's worker_pipeline does for one mini-batch
 * with its default options: mp_map_coarse, mp_map_finish and the output step, map.c:547-870,
 * 1143-1166; no --accuracy, HMM queries, --prescreen, --cand-table or --gpu-prefilter). The mapper
 * holds the genome (a copy of the base codes 0..4, contigs one after another, and their names) and its
 * k-mer block index on the device (mp_idx_build's options: min_aa_len, kmer, mod_bit, bbit; the tables
 * must be set up first, ns_make_tables). nsh_map_batch maps n_q proteins (ASCII) and returns the
 * records mp_write_output writes for them, in query order, in one malloc'd NUL-terminated buffer
 * (*out, *out_len bytes; free() it): PAF, or GFF3/GTF by opt->flag (MP_F_GFF 0x8, MP_F_NO_PAF 0x10,
 * MP_F_GTF 0x20, MP_F_NO_CS 0x200); the "##gff-version 3" line that mp_map_file prints first is the
 * caller's, and unmapped proteins write nothing (MP_F_SHOW_UNMAP is not supported). *id is the
 * output step's running record id (pipeline_t.id), advanced by the records written. stats (24
 * entries, may be NULL): [0] coarse chains, [1] refine candidates, [2] refine anchor pairs, [3] DP
 * candidates, [4] proteins with a record, [5] records, [6] candidates offered to the prescreen,
 * [7] candidates it dropped, [8..13] microseconds since the call began at the end of: the coarse
 * chains, the region selection, the refine set-up, the refinement, the align stage, the
 * statistics and records; [14] rejected candidates kept by the audit, [15] of those, the ones that
 * became records; [16] microseconds at the end of the prescreen. 1 ok, 0 error (reason in
 * ns_gpu_error).
 *
 * The learned prescreen (no reference counterpart; the reference's --gpu-prefilter sits at the same
 * point, map.c:1114-1117) runs between the coarse region selection and the refinement when
 * opt->prescreen is set: each candidate's 64 features (the device feature kernel over the region's
 * window, at most 24,000 bases, and the protein's first 4,096 residues; slots 46-59 replaced by the
 * chain's own: score over length and over the query's best, ungapped score, anchors, query
 * coverage and ends, span, primary flag, rank, candidate count, score over its primary's, gap to
 * the best), the network's logit, and the cut at ps_tau; a ps_audit fraction of the rejected
 * (drawn by a hash of ps_seed and the candidate) is kept anyway, and stats[15] counts the audited
 * ones that became records. cand_feat / cand_label (both optional, room for cand_cap candidates)
 * receive every offered candidate's features and whether it became a record, in query order: the
 * training data. */
typedef struct {
	nsh_align_opt_t dp;                   /* the DP options, and max_ext, max_intron, kmer2 */
	int32_t max_occ, max_gap, bw, max_chn_max_skip, max_chn_iter, min_chn_cnt, min_chn_sc;
	float chn_coef_log;
	int32_t max_ava, best_n, out_n, mask_len;
	float pri_ratio, mask_level, out_sim, out_cov;
	uint32_t flag;                        /* output flags, as above */
	int32_t gff_delim;
	const char *gff_prefix;               /* NULL: "PL" */
	int32_t n_threads;                    /* host threads of the host steps */
	struct nsh_mlp_s *prescreen;          /* the learned candidate prescreen (nsh_mlp_*); NULL: none */
	float ps_tau, ps_audit;               /* keep logit >= ps_tau, and a ps_audit fraction of the rest */
	uint64_t ps_seed;
	float *cand_feat;                     /* optional out: 64 features per candidate */
	uint8_t *cand_label;                  /* optional out: 1 when the candidate became a record */
	int64_t cand_cap;                     /* room in cand_feat / cand_label, in candidates */
} nsh_map_opt_t;
typedef struct nsh_mapper_s nsh_mapper_t;
void nsh_map_opt_default(nsh_map_opt_t *o);   /* mp_mapopt_init (options.c:42-138), 16 host threads */
nsh_mapper_t *nsh_mapper_init(int32_t n_ctg, const char *const *ctg_name, const uint8_t *nt4, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit);
int nsh_map_batch(nsh_mapper_t *m, int32_t n_q, const char *const *qname, const char *const *seq, const int32_t *qlen,
		const nsh_map_opt_t *opt, int64_t *id, char **out, int64_t *out_len, int64_t *stats);
void nsh_mapper_destroy(nsh_mapper_t *m);

/* Batched global spliced DP with CIGAR (no reference counterpart; the batch form of
 * ns_global_gs16b / ns_global_gs32b / ns_global_gs_auto with NS_F_CIGAR): n problems with one option
 * set (opt->flag must hold NS_F_CIGAR and no extension flag; opt->pssm applies to every problem, so it
 * is normally NULL), ns/as concatenated with offsets, ss optional (ns offsets). mode 0: int16
 * (gs16b), 1: int32 (gs32b), 2: per problem as ns_global_gs_auto. All problems go to the engine as one
 * batch. Results in input order: score[k], n_cigar[k]; *cigar receives one malloc'd array holding every
 * CIGAR (len << 4 | op) one after another (free() it; NULL when no problem has one). 1 ok, 0 error. */
int nsh_global_batch(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *ss, const ns_opt_t *opt, int32_t mode,
		int32_t *score, int32_t *n_cigar, uint32_t **cigar);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): nsh_ext_plan_create + run +
 * results in one call that calls no hipMalloc/hipFree. Arguments as nsh_ext_plan_create (ss and a
 * PSSM are not supported), results as nsh_ext_plan_results (input order); the work is queued on
 * `stream` (a hipStream_t; NULL = the null stream) and waited for before returning. Device memory
 * comes from one fixed 128-MB array per GPU held until the process ends; limits: 16,384 problems,
 * DNA <= 262,143 nt and padded protein <= 2,048 per problem, and the inputs' device footprint
 * (about 5 B per DNA base + 2 B per residue + 144 B per problem) within 128 MB; beyond them it
 * returns 0 with the reason in ns_gpu_error(). Calls on one GPU share the array: they are
 * serialised by the library, and two processes' calls do not share it. 1 ok, 0 error. */
int nsh_ext_run_noalloc(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl,
		const char *as, const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt,
		nsh_ext_result_t *out, void *stream);

/* Profile-HMM engine behind mp_hmm_dp_score_gpu*: a profile resident in HBM (emissions M x 20 in
 * HMMER column order, best residue at +4, transitions M x 7 m->m m->i m->d i->m i->i d->m d->d, as
 * mp_hmm_read1 stores them, hmm.c:112-195). Scores equal mp_hmm_dp_score (hmm.c:252-291). M <= 1024. */
typedef struct nsh_hmm_s nsh_hmm_t;
nsh_hmm_t *nsh_hmm_create(int32_t M, const int16_t *emit, const int16_t *trans);
int nsh_hmm_score_batch(nsh_hmm_t *h, int32_t n, const uint8_t *nt, int64_t l_nt, const int32_t *off,
		const int32_t *len, int32_t *score);
float nsh_hmm_last_ms(nsh_hmm_t *h);   /* device time of the last batch's Viterbi kernel (ms) */
void nsh_hmm_destroy(nsh_hmm_t *h);

/* Best ungapped PSSM segment of item k: ps_ungapped(pssm, L, aa + orf_off[k], orf_len[k], diag[k])
 * (prescreen.c:334-344) -> score[k], query interval [qs[k], qe[k]). pssm: L x 22 int8 rows. */
int nsh_pssm_ungapped_batch(const int8_t *pssm, int32_t L, const uint8_t *aa, int64_t l_aa, int32_t n,
		const int64_t *orf_off, const int32_t *orf_len, const int32_t *diag, int32_t *score, int32_t *qs, int32_t *qe);

/* PSSM diagonal evidence of refine chains: ev[c] = mp_chain_pssm_ev(nt + nt_off[c], l_nt[c],
 * pssm + pssm_off[c], qlen[c], kmer, n_a[c], a + a_off[c], min_seg, xdrop) (map.c:123-139, with
 * mp_diag_seg, map.c:97-120). nt: region bases (codes 0..4); pssm: qlen x 22 int8 rows; anchors
 * x << 32 | q (x the k-mer's last base in the region, q its last residue; q + 1 - kmer must lie in
 * [0, qlen]). seg_sc / seg_e (both or neither, one per anchor): the anchor's segment score and
 * protein end, INT32_MIN / -1 when it repeats a recorded segment. 1 on success, 0 on error. */
int nsh_chain_pssm_ev_batch(int32_t n, const uint8_t *nt, int64_t l_nt_all, const int64_t *nt_off,
		const int64_t *l_nt, const int8_t *pssm, int64_t l_pssm, const int64_t *pssm_off, const int32_t *qlen,
		const uint64_t *a, int64_t n_a_all, const int64_t *a_off, const int32_t *n_a, int32_t kmer,
		int32_t min_seg, int32_t xdrop, int32_t *ev, int32_t *seg_sc, int32_t *seg_e);

/* Six-frame ORF walk of ps_orf_worker + ps_add_orf (prescreen.c:87-132) over one oriented
 * sequence (base codes 0..4; the reverse strand is passed reverse-complemented, as mp_ntseq_get
 * returns it): every stop- and N-free run of >= min_orf (>= 1) codons in each frame, in the
 * reference's order. orf_st[k] is the run's first base, orf_len[k] its codon count, and aa holds
 * the runs' residues back to back (*n_aa_total of them). At most max_orf runs and max_aa
 * residues are written. Returns the number of runs, or -1 on error. */
int64_t nsh_orf_scan(const uint8_t *nt, int64_t len, int32_t min_orf, int64_t max_orf, int64_t *orf_st,
		int32_t *orf_len, int64_t max_aa, uint8_t *aa, int64_t *n_aa_total);

/* Banded gapped stage of the prescreen search: sw[k] = ps_gapped(pssm, L, aa + orf_off[k],
 * orf_len[k], diag[k], band, go, ge) (prescreen.c:346-382), the best local affine score of the
 * profile (L x 22 int8 rows) against ORF k (residue codes < 22) within |t - i - diag| <= band, or
 * the full matrix when band <= 0. Requires 0 <= ge <= go <= 65535 and 2 band + 1 (or, for
 * band <= 0, the ORF length) <= 1024. 1 on success, 0 on error. */
int nsh_ps_gapped_batch(const int8_t *pssm, int32_t L, const uint8_t *aa, int64_t l_aa, int32_t n,
		const int64_t *orf_off, const int32_t *orf_len, const int32_t *diag, int32_t band, int32_t go, int32_t ge,
		int32_t *sw);

/* The profile prescreen search (prescreen.c:133-486): an ORF database and spaced k-mer index of a
 * genome built once on the device (mp_prescreen_build: contigs of base codes 0..4, forward
 * strand, at nt + ctg_off[c]; kmer 5..7; ORFs of >= max(min_orf, span) aa on both strands), then
 * per profile the reference worker's candidates (ps_query_worker): pssm L x 22 int8 rows, thr the
 * k-mer score threshold (mp_ps_kmer_thr), kmer_max, min_ungapped, max_seqs, band, go, ge and
 * max_evalue as mp_psopt_t holds them. nsh_ps_cand_t has mp_cand1_t's layout (protlocus.h:73-80);
 * the candidates come in the worker's order, before mp_cand_tab_set merges them. The search
 * returns how many there are (only max_out are written), or -1 on error.
 * nsh_prescreen_size: 0 ORFs, 1 residues, 2 index entries. */
typedef struct nsh_prescreen_s nsh_prescreen_t;
typedef struct {
	uint32_t vid;               /* contig << 1 | reverse */
	int64_t vs, ve;             /* the ORF's oriented span */
	double ev;                  /* E-value of its banded gapped score */
	int32_t aqs, aqe, asc;      /* best ungapped diagonal segment: query interval and score */
	int64_t avs;                /* oriented start of the codon facing query position aqs */
} nsh_ps_cand_t;
nsh_prescreen_t *nsh_prescreen_build(int32_t n_ctg, const uint8_t *nt, const int64_t *ctg_off,
		const int64_t *ctg_len, int32_t kmer, int32_t min_orf);
int64_t nsh_prescreen_size(const nsh_prescreen_t *ps, int32_t what);
int32_t nsh_prescreen_search(nsh_prescreen_t *ps, const int8_t *pssm, int32_t L, int32_t thr, int32_t kmer_max,
		int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge, double max_evalue,
		nsh_ps_cand_t *out, int32_t max_out);
/* The search of n_prof profiles at once (mp_prescreen_batch, prescreen.c:488-501, which runs
 * ps_query_worker for every profile): pssm holds the profiles' L[i] x 22 rows one after another.
 * The candidates come profile after profile, each profile's list as nsh_prescreen_search returns
 * it; n_res[i] receives profile i's count. Returns the total (only max_out are written), or -1. */
int64_t nsh_prescreen_search_batch(nsh_prescreen_t *ps, int32_t n_prof, const int8_t *pssm, const int32_t *L,
		int32_t thr, int32_t kmer_max, int32_t min_ungapped, int32_t max_seqs, int32_t band, int32_t go, int32_t ge,
		double max_evalue, nsh_ps_cand_t *out, int64_t max_out, int32_t *n_res);
/* Counters of the last search (the figures mp_prescreen_batch logs, prescreen.c:503-525): st[0]
 * query windows, st[1] k-mers generated, st[2] index hits, st[3] double-hit diagonals, st[4] ORFs
 * whose best diagonal reaches min_ungapped, st[5] aligned (banded gapped), st[6] kept. 1 or 0. */
int nsh_prescreen_stats(nsh_prescreen_t *ps, int64_t *st);
void nsh_prescreen_destroy(nsh_prescreen_t *ps);

/* The k-mer block index of a genome (mp_idx_build without the file read: mp_idx_boff,
 * mp_sketch_nt4 per contig strand and build_bidx, index.c:14-29, 52-97, sketch.c:108-169), built on
 * the device: contigs of base codes 0..4 at nt + ctg_off[c]; k-mers of 1..7 residues in the 13-letter
 * alphabet (ns_tab_codon13) over clean ORFs of >= min_aa_len codons, kept when their low mod_bit
 * hash bits are 0; blocks of 2^bbit bases. nsh_index_size: 0 entries (n_kb), 1 buckets
 * (2^(4 kmer - mod_bit)), 2 blocks (n_block). nsh_index_copy: the bucket starts (mp_idx_t::ki, one
 * per bucket), the entries (kb) and the block offsets (bo, 2 n_ctg + 1) into the arrays that are not
 * NULL; 1 on success. The tables and codes must be set up (ns_make_tables) before the build. */
typedef struct nsh_index_s nsh_index_t;
nsh_index_t *nsh_index_build(int32_t n_ctg, const uint8_t *nt, const int64_t *ctg_off, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit);
/* the same index from the reference's 4-bit genome as mp_ntseq_read stores it (mp_ntdb_t.seq:
 * base i in the low nibble of byte i/2 when i is even, the high one when odd; ctg_off in bases):
 * half the bytes cross PCIe, and the device unpacks them */
nsh_index_t *nsh_index_build_nt4(int32_t n_ctg, const uint8_t *seq4, const int64_t *ctg_off, const int64_t *ctg_len,
		int32_t min_aa_len, int32_t kmer, int32_t mod_bit, int32_t bbit);
int64_t nsh_index_size(const nsh_index_t *ix, int32_t what);
int nsh_index_copy(const nsh_index_t *ix, int64_t *ki, uint32_t *kb, uint32_t *bo);
void nsh_index_destroy(nsh_index_t *ix);
/* The anchors of n_q queries (ASCII residues at seq + q_off[q], q_len[q] of them) against the
 * index, as mp_map_coarse collects them with the default options (map.c:583-597): the protein
 * sketch (mp_sketch_prot, sketch.c:18-38), the repeat cutoff (mp_cal_max_occ, map.c:304-318)
 * capped at max_occ, the bucket entries of the seeds under it (mp_collect_anchors, map.c:336-395,
 * without the --accuracy repeat rescue), each query's anchors sorted. An anchor is
 * block << 32 | query position. Query q's anchors are anchors[a_off[q] .. a_off[q+1]) (a_off has
 * n_q + 1 entries); returns the total (only max_anchors are written), or -1 on error. */
int64_t nsh_index_anchors(const nsh_index_t *ix, int32_t n_q, const char *seq, const int64_t *q_off,
		const int32_t *q_len, int32_t max_occ, int64_t *a_off, uint64_t *anchors, int64_t max_anchors);

/* Anchor chaining of n_q queries (mp_chain, chain.c:7-219, each query's call as mp_map_coarse
 * makes it, map.c:611-632): query q's sorted anchors are anchors[a_off[q] .. a_off[q+1]); the
 * parameters are mp_chain's. n_u[q] receives its chain count; its chains (score << 32 | anchors)
 * go to u + a_off[q] and their anchors, chain after chain in mp_chain's order, to
 * a_out + a_off[q] (u and a_out have a_off[n_q] slots). Returns the total chain count, or -1. */
int64_t nsh_chain_batch(int32_t n_q, const int64_t *a_off, const uint64_t *anchors, int32_t max_dist_x,
		int32_t max_dist_y, int32_t bw, int32_t max_skip, int32_t max_iter, int32_t min_cnt, int32_t min_sc,
		float chn_coef_log, int32_t is_spliced, int32_t kmer, int32_t bbit, int32_t *n_u, uint64_t *u, uint64_t *a_out);

/* mp_map_coarse's index -> chain slice for n_q queries (map.c:583-633 with the default options,
 * no --accuracy vid chaining): the anchors as nsh_index_anchors collects them, then, when
 * pre_chain and is_spliced are set, the pre-chain (mp_chain with max_dist_x = max_dist_y = bw =
 * 1 << bbit, min_cnt 2, min_sc 0) and each query's kept anchors sorted (radix_sort_mp64), then
 * mp_chain with these options; kmer and bbit are the index's. Everything stays on the device.
 * Query q's chains (score << 32 | anchors) are u[u_off[q] .. u_off[q+1]) and their anchors, chain
 * after chain, anchors[c_off[q] .. c_off[q+1]); u_off and c_off (n_q + 1 entries) are always
 * written, at most max_u chains and max_anchors anchors. Returns the total chain count, or -1. */
typedef struct {
	int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_sc;   /* max_intron, max_gap, ... */
	float chn_coef_log;
	int32_t is_spliced, pre_chain;
} nsh_chain_opt_t;
int64_t nsh_index_chain(const nsh_index_t *ix, int32_t n_q, const char *seq, const int64_t *q_off,
		const int32_t *q_len, int32_t max_occ, const nsh_chain_opt_t *opt, int64_t *u_off, uint64_t *u,
		int64_t max_u, int64_t *c_off, uint64_t *anchors, int64_t max_anchors);

/* Learned chain-candidate prescreen: MLP 64 -> 256 -> 256 -> 1 (ReLU, logistic output), BF16
 * MFMA kernels with FP32 accumulation, FP32 Adam masters. Parameter vector (FP32, in order):
 * W1[64][256], b1[256], W2[256][256], b2[256], w3[256], b3. No reference counterpart. */
typedef struct nsh_mlp_s nsh_mlp_t;
int32_t nsh_mlp_param_count(void);
nsh_mlp_t *nsh_mlp_create(const float *params);
/* one step on B samples (B % 32 == 0): forward, gradients (mean BCE), Adam when lr > 0 */
int nsh_mlp_train_step(nsh_mlp_t *h, const float *X, const float *t, int32_t B, float lr, float *loss);
int nsh_mlp_get_params(nsh_mlp_t *h, float *params, float *grads);
/* logits of n candidates; X: n x 64 FP32 features on the host (x_on_device = 0) or device */
int nsh_mlp_infer(nsh_mlp_t *h, const float *X, int64_t n, int32_t x_on_device, float *logit);
float nsh_mlp_last_ms(nsh_mlp_t *h);   /* device time of the last inference (ms) */
/* standardise inputs on the device, (x - mu) / sd per feature, in training and inference */
int nsh_mlp_set_norm(nsh_mlp_t *h, const float *mu, const float *sd);
void nsh_mlp_destroy(nsh_mlp_t *h);

/* the same features from raw host strings (proteins ASCII, DNA ASCII or 0..4), segment scores
 * with BLOSUM62 (plans use their own matrix); out: n x 64 */
int nsh_features(int32_t n, const char *as, const int64_t *as_off, const int32_t *al, int64_t l_as,
		const char *ns, const int64_t *ns_off, const int32_t *nl, int64_t l_ns, float *out);

/* ======================================================================================
 * Alignment output: statistics, CDS features and the PAF / GFF3 / GTF records of one
 * alignment, byte-identical to the reference's writers. Host code.
 * ====================================================================================== */
typedef struct {                /* one alignment: mp_reg1_t with its mp_extra_t (protlocus.h:205-254) */
	const char *qname, *aa;     /* protein name and sequence (ASCII) */
	int32_t qlen, qs, qe;       /* protein length, aligned interval [qs, qe) */
	const char *ctg_name;       /* contig name and length */
	int64_t ctg_len;
	int32_t rev;                /* strand: 0 forward, 1 reverse (mp_reg1_t.vid & 1) */
	int64_t vs, ve;             /* aligned target interval in oriented coordinates */
	int32_t dp_score;           /* DP score (mp_extra_t.dp_score) */
	int32_t profile_score;      /* HMM Viterbi score, INT32_MIN when none (PAF ps:i:) */
	int32_t rank_sc;            /* rerank score (PAF ra:i: with MP_F_ACCURACY and MP_ACC_RERANK) */
	int32_t chn_sc, chn_sc_ungap, cnt;   /* chain fields, printed when the details were dropped */
	const uint32_t *cigar;      /* len << 4 | op, as ns_rst_t.cigar */
	int32_t n_cigar;
} nsh_aln_t;

typedef struct {                /* what mp_extra_stop / mp_extra_start / mp_extra_cal fill into mp_extra_t */
	int32_t dp_max, blen, n_fs, n_stop, n_iden, n_plus, dist_stop, dist_start;
	int32_t n_exon, n_feat;
} nsh_aln_stats_t;

typedef struct {                /* one CDS (type 0) or stop_codon (type 1) feature; layout of mp_feat_t (protlocus.h:230-237) */
	int64_t vs, ve;
	int32_t qs, qe;
	int16_t type, phase;
	int32_t n_fs, n_stop;
	int32_t score, n_iden, blen;
	char donor[2], acceptor[2];
} nsh_feat_t;

/* Statistics and features of an alignment: mp_extra_stop, mp_extra_start and mp_extra_cal
 * (align.c:148-306) as align.c:608-612 calls them. nt: oriented target bases (codes 0..4) of
 * [as, ae), covering [vs, ve), searched for the nearest stop and start codons. pssm: the PSSM
 * row of protein position qs, rows of opt->asize (HMM mode, mp_pssm_at(c, qs)), or NULL to score
 * with opt->sc. Returns the number
 * of features (>= 1); 0 when the CIGAR does not span [qs, qe) x [vs, ve), where the reference
 * drops the alignment's details (align.c:258-268); -1 on bad arguments or m_feat too small. */
int32_t nsh_aln_stats(const nsh_aln_t *a, const uint8_t *nt, int64_t as, int64_t ae, const ns_opt_t *opt,
		const int8_t *pssm, nsh_aln_stats_t *st, nsh_feat_t *feat, int32_t m_feat);
/* The records mp_write_output (format.c:457-477) writes for the alignment: flag and acc_flags
 * are mp_mapopt_t.flag and .acc_flags (MP_F_GFF 0x8, MP_F_NO_PAF 0x10, MP_F_GTF 0x20,
 * MP_F_NO_CS 0x200, MP_F_ACCURACY 0x400; MP_ACC_RERANK 0x8), gff_prefix (NULL: "PL"), id and
 * gff_delim as the reference passes them, hit_idx the hit's rank. st and feat from
 * nsh_aln_stats, st == NULL when it returned 0. Writes at most size - 1 bytes and a NUL; returns
 * the full length, or -1 (MP_F_SHOW_RESIDUE and MP_F_SHOW_TRANS are not supported). */
int64_t nsh_write_output(char *out, int64_t size, const nsh_aln_t *a, const nsh_aln_stats_t *st,
		const nsh_feat_t *feat, const uint8_t *nt, int64_t as, int32_t flag, int32_t acc_flags,
		const char *gff_prefix, int64_t id, int32_t gff_delim, int32_t hit_idx);

int nsh_device_synchronize(void);
const char *nsh_build_info(void);

// Measurement aid, no counterpart in the reference: reads `bytes` (a multiple of 64 KiB) of device
// memory at src through the scalar data cache, one wave per 64 KiB, and writes one folded word per
// wave to out (device memory, bytes / 65536 words); synchronous. A read of known size on the path
// the extension DP's row records take, for scaling hardware counter readings. 1 on success.
int nsh_diag_scalar_read(const void *src, int64_t bytes, uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif
