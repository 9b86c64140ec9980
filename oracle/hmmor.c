/*
 * hmmor.c -- oracle restatement of the profile-HMM reader and Viterbi (hmm.c), the
 * gapless X-drop candidate score (nasw-cuda.cu) and the PSSM ungapped diagonal scan
 * (prescreen.c). Test infrastructure only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <math.h>
#include <ctype.h>
#include <limits.h>
#include "nsor.h"

/* HMMER3 emission column order and the HMMER amino-acid background (hmm.c:26-31) */
static const char or_hmm_alpha[] = "ACDEFGHIKLMNPQRSTVWY";
static const double or_hmm_bg[20] = {
	0.0787945, 0.0151600, 0.0535222, 0.0668298, 0.0397062, 0.0695071, 0.0229198, 0.0590092, 0.0594422, 0.0963728,
	0.0237718, 0.0414386, 0.0482904, 0.0395639, 0.0540978, 0.0683364, 0.0540687, 0.0673417, 0.0114135, 0.0304133
};

/* hmm.c:46-58: "*" and unparsable -> -32768; else lrint(-2x) clamped to +-32767 */
static int or_parse_score(const char *s)
{
	char *end;
	double x;
	if (strcmp(s, "*") == 0) return -32768;
	errno = 0;
	x = strtod(s, &end);
	if (errno || end == s || *end) return -32768;
	x = -2.0 * x;
	if (x < -32767.) x = -32767.;
	if (x > 32767.) x = 32767.;
	return (int)lrint(x);
}

/* hmm.c:69-81: half-bit log-odds vs background, clamped to +-64 */
static int8_t or_pssm_score(const char *s, double bg)
{
	char *end;
	double x, lo;
	if (strcmp(s, "*") == 0) return -64;
	errno = 0;
	x = strtod(s, &end);
	if (errno || end == s || *end) return -64;
	lo = 2.0 * (-x - log(bg)) / log(2.0);
	if (lo < -64.) lo = -64.;
	if (lo > 64.) lo = 64.;
	return (int8_t)lrint(lo);
}

static int or_split(char *line, char **tok, int max_tok)
{
	int n = 0;
	char *p = strtok(line, " \t\r\n");
	while (p && n < max_tok) tok[n++] = p, p = strtok(0, " \t\r\n");
	return n;
}

static int or_is_pos_line(const char *line)
{
	char *end;
	(void)strtol(line, &end, 10);
	return end != line && isspace((unsigned char)*end);
}

/* line reader over an in-memory text */
static int or_getline(const char **p, char *buf, int cap)
{
	int n = 0;
	if (!**p) return 0;
	while (**p && **p != '\n' && n < cap - 2) buf[n++] = *(*p)++;
	if (**p == '\n') buf[n++] = *(*p)++;
	buf[n] = 0;
	return 1;
}

/* hmm.c:112-195 (mp_hmm_read1), reading from a string; returns 1 on success, 0 at end, -1 on error */
int or_hmm_parse(const char *text, or_hmm_t *h, const char **next)
{
	char line[8192], copy[8192], *tok[64];
	const char *p = text;
	int in_hmm = 0, have_header = 0, pos = 0, i, n;
	memset(h, 0, sizeof(*h));
	while (or_getline(&p, line, sizeof(line))) {
		char *data = line;
		while (*data && isspace((unsigned char)*data)) ++data;
		if (strncmp(line, "LENG", 4) == 0) {
			if (sscanf(line + 4, "%d", &h->len) != 1 || h->len <= 0) goto fail;
			free(h->emit), free(h->trans), free(h->pssm);
			h->emit = (int16_t*)calloc((size_t)h->len * 20, sizeof(int16_t));
			h->trans = (int16_t*)calloc((size_t)h->len * 7, sizeof(int16_t));
			h->pssm = (int8_t*)calloc((size_t)h->len * 22, 1);
		} else if (strncmp(line, "HMM", 3) == 0 && isspace((unsigned char)line[3])) {
			in_hmm = 1, have_header = 1;
		} else if (in_hmm && line[0] == '/') {
			break;
		} else if (in_hmm && or_is_pos_line(data)) {
			if (*data == '0') { /* COMPO line: skip it and the two that follow */
				if (!or_getline(&p, line, sizeof(line)) || !or_getline(&p, line, sizeof(line))) break;
				continue;
			}
			if (pos >= h->len) goto fail;
			strncpy(copy, data, sizeof(copy) - 1), copy[sizeof(copy) - 1] = 0;
			n = or_split(copy, tok, 64);
			if (n < 21) goto fail;
			{
				int best = -32768;
				int8_t *row = h->pssm + (size_t)pos * 22;
				for (i = 0; i < 20; ++i) {
					h->emit[pos * 20 + i] = (int16_t)or_parse_score(tok[i + 1]);
					if (h->emit[pos * 20 + i] > best) best = h->emit[pos * 20 + i];
					row[or_aa20[(uint8_t)or_hmm_alpha[i]]] = or_pssm_score(tok[i + 1], or_hmm_bg[i]);
				}
				row[or_aa20['*']] = -23;
				row[or_aa20['X']] = -1;
				for (i = 0; i < 20; ++i) h->emit[pos * 20 + i] = (int16_t)(h->emit[pos * 20 + i] - best + 4);
			}
			++pos;
			if (!or_getline(&p, line, sizeof(line)) || !or_getline(&p, line, sizeof(line))) break;
			strncpy(copy, line, sizeof(copy) - 1), copy[sizeof(copy) - 1] = 0;
			n = or_split(copy, tok, 16);
			if (n >= 7)
				for (i = 0; i < 7; ++i) h->trans[(pos - 1) * 7 + i] = (int16_t)or_parse_score(tok[i]);
		}
	}
	if (next) *next = p;
	if (!have_header && !h->emit && pos == 0) { or_hmm_free(h); return 0; }
	if (!have_header || !h->emit || pos != h->len) goto fail;
	return 1;
fail:
	or_hmm_free(h);
	if (next) *next = p;
	return -1;
}

void or_hmm_free(or_hmm_t *h)
{
	free(h->emit), free(h->trans), free(h->pssm);
	memset(h, 0, sizeof(*h));
}

/* hmm.c:225-233: aa20 index -> HMMER column (-1 for stop / X) */
const int8_t *or_hmm_aa20_to_alpha(void)
{
	static int8_t t[22];
	int i;
	for (i = 0; i < 22; ++i) t[i] = -1;
	for (i = 0; i < 20; ++i) t[or_aa20[(uint8_t)or_hmm_alpha[i]]] = (int8_t)i;
	return t;
}

/* hmm.c:252-291: local Viterbi over frame-0 codons of the region, int32 with clamp at 0 */
int32_t or_hmm_viterbi(const or_hmm_t *h, const uint8_t *nt, int32_t nlen)
{
	int32_t M = h->len, i, j, rows, best = 0;
	int32_t *pm, *pi, *pd, *cm, *ci, *cd, *x;
	const int8_t *a2a = or_hmm_aa20_to_alpha();
	if (!nt || nlen < 3 || M <= 0) return 0;
	rows = nlen / 3;
	pm = (int32_t*)calloc((size_t)M + 1, 4), pi = (int32_t*)calloc((size_t)M + 1, 4), pd = (int32_t*)calloc((size_t)M + 1, 4);
	cm = (int32_t*)calloc((size_t)M + 1, 4), ci = (int32_t*)calloc((size_t)M + 1, 4), cd = (int32_t*)calloc((size_t)M + 1, 4);
	for (i = 0; i < rows; ++i) {
		const uint8_t *c = nt + 3 * i;
		int32_t r = (c[0] > 3 || c[1] > 3 || c[2] > 3)? 20 : or_codon[c[0] << 4 | c[1] << 2 | c[2]];
		int32_t e = r < 20? a2a[r] : -1;
		for (j = 1; j <= M; ++j) {
			const int16_t *t = h->trans + (size_t)(j - 1) * 7;
			int32_t from = pm[j-1] + t[0], v;
			if (pi[j-1] + t[3] > from) from = pi[j-1] + t[3];
			if (pd[j-1] + t[5] > from) from = pd[j-1] + t[5];
			v = from + (e >= 0? h->emit[(size_t)(j - 1) * 20 + e] : 0);
			cm[j] = v < 0? 0 : v;
			v = pm[j] + t[1] > pi[j] + t[4]? pm[j] + t[1] : pi[j] + t[4];
			ci[j] = v < 0? 0 : v;
			v = pm[j-1] + t[2] > pd[j-1] + t[6]? pm[j-1] + t[2] : pd[j-1] + t[6];
			cd[j] = v < 0? 0 : v;
			if (cm[j] > best) best = cm[j];
			if (ci[j] > best) best = ci[j];
		}
		x = pm, pm = cm, cm = x;
		x = pi, pi = ci, ci = x;
		x = pd, pd = cd, cd = x;
	}
	free(pm), free(pi), free(pd), free(cm), free(ci), free(cd);
	return best;
}

/* ------------------------------------------------------------------------------------ */
/* gapless X-drop score of one candidate: nasw-cuda.cu:1151-1295                          */
/* ------------------------------------------------------------------------------------ */
static uint8_t or_gl_base(const uint8_t *pk, int64_t off, int64_t len, int32_t rev, int64_t pos)
{
	int64_t a;
	uint8_t b;
	if (pos < 0 || pos >= len) return 4;
	a = off + (rev? len - 1 - pos : pos);
	b = (uint8_t)(pk[a >> 1] >> ((a & 1) * 4) & 0xf);
	return rev && b < 4? (uint8_t)(3 - b) : b;
}

static uint8_t or_gl_aa(const uint8_t *pk, const uint8_t *codon, int64_t off, int64_t len, int32_t rev, int64_t end)
{
	uint8_t b0 = or_gl_base(pk, off, len, rev, end - 2), b1 = or_gl_base(pk, off, len, rev, end - 1), b2 = or_gl_base(pk, off, len, rev, end);
	return b0 < 4 && b1 < 4 && b2 < 4? codon[b0 << 4 | b1 << 2 | b2] : 20;
}

int32_t or_gapless_score(const uint8_t *pk, const uint8_t *q, const ns_gpu_gapless_candidate_t *c,
		const ns_gpu_gapless_anchor_t *anchors, int32_t block_width, int32_t kmer, int32_t xdrop,
		const int8_t *mat, int32_t asize, const uint8_t *codon)
{
	int32_t na = c->anchor_count < 32? c->anchor_count : 32, a, n_seg = 0, total = 0, any = 0, i, j;
	int32_t seg_qs[32], seg_qe[32], seg_sc[32];
	const uint8_t *qq = q + c->query_offset;
	for (a = 0; a < na; ++a) {
		const ns_gpu_gapless_anchor_t *an = &anchors[c->anchor_offset + a];
		int32_t best = INT32_MIN, best_q = -1, d, s, sc, qs, qe;
		int64_t best_t = INT64_MAX;
		if (an->query_end >= kmer - 1 && an->query_end < c->query_length) {
			for (d = 0; d < block_width; ++d) { /* best k-mer seed; ties -> smallest target end */
				int64_t te = an->target_block + d;
				if (te - (int64_t)(kmer - 1) * 3 - 2 < 0 || te >= c->target_length) continue;
				for (s = 0, sc = 0; s < kmer; ++s) {
					uint8_t ta = or_gl_aa(pk, codon, c->target_offset, c->target_length, c->reverse, te - (int64_t)(kmer - 1 - s) * 3);
					sc += mat[ta * asize + qq[an->query_end - kmer + 1 + s]];
				}
				if (sc > best || (sc == best && te < best_t)) best = sc, best_t = te, best_q = an->query_end;
			}
		}
		if (best_q < 0) continue; /* no usable seed: this anchor contributes no segment */
		{
			int32_t lr = 0, lb = 0, ll = 0, rr = 0, rb = 0, rl = 0, qp, st;
			for (qp = best_q - kmer, st = 1; qp >= 0; --qp, ++st) {
				int64_t te = best_t - (int64_t)(kmer - 1 + st) * 3;
				if (te - 2 < 0) break;
				lr += mat[or_gl_aa(pk, codon, c->target_offset, c->target_length, c->reverse, te) * asize + qq[qp]];
				if (lr > lb) lb = lr, ll = st;
				if (lb - lr > xdrop) break;
			}
			for (qp = best_q + 1, st = 1; qp < c->query_length; ++qp, ++st) {
				int64_t te = best_t + (int64_t)st * 3;
				if (te >= c->target_length) break;
				rr += mat[or_gl_aa(pk, codon, c->target_offset, c->target_length, c->reverse, te) * asize + qq[qp]];
				if (rr > rb) rb = rr, rl = st;
				if (rb - rr > xdrop) break;
			}
			sc = best + lb + rb, qs = best_q - (kmer - 1) - ll, qe = best_q + rl;
		}
		/* insertion into the score-descending segment list (stable for ties: later goes after) */
		any = 1;
		for (j = n_seg; j > 0 && seg_sc[j - 1] < sc; --j)
			seg_qs[j] = seg_qs[j - 1], seg_qe[j] = seg_qe[j - 1], seg_sc[j] = seg_sc[j - 1];
		seg_qs[j] = qs, seg_qe[j] = qe, seg_sc[j] = sc, ++n_seg;
	}
	for (i = 0; i < n_seg; ++i) {
		int32_t len = seg_qe[i] - seg_qs[i] + 1, ov = 0;
		if (i > 0 && seg_sc[i] < xdrop) break;
		for (j = 0; j < i; ++j) {
			int32_t s = seg_qs[i] > seg_qs[j]? seg_qs[i] : seg_qs[j], e = seg_qe[i] < seg_qe[j]? seg_qe[i] : seg_qe[j];
			if (seg_sc[j] != INT32_MIN && e >= s) ov += e - s + 1;
		}
		if (ov >= len) { seg_sc[i] = INT32_MIN; continue; }
		if (ov > 0) total += (int32_t)((int64_t)seg_sc[i] * (len - ov) / len);
		else total += seg_sc[i];
	}
	return any? total : INT32_MIN;
}

/* prescreen.c:334-344 */
int32_t or_pssm_ungapped(const int8_t *pssm, int32_t L, const uint8_t *orf, int32_t n, int32_t diag, int32_t *qs, int32_t *qe)
{
	int32_t i, i0 = diag < 0? -diag : 0, i1 = n - diag < L? n - diag : L, sc = 0, best = 0, st = i0;
	*qs = *qe = i0;
	for (i = i0; i < i1; ++i) {
		sc += pssm[(int64_t)i * 22 + orf[i + diag]];
		if (sc < 0) sc = 0, st = i + 1;
		else if (sc > best) best = sc, *qs = st, *qe = i + 1;
	}
	return best;
}
