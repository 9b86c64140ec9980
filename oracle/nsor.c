/*
 * nsor.c -- oracle restatement of the reference's CPU DP semantics (test infrastructure:
 * used only by tests/, bench.py's CPU-baseline leg and __graft_entry__.smoke()).
 *
 * Written from the reference's recurrences; the striped global path is emulated lane by
 * lane so that its trace bits (which depend on the SIMD evaluation order) come out exactly.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <limits.h>
#include "nsor.h"

/* ------------------------------------------------------------------------------------ */
/* tables: nasw-tab.c:6-107. The 33 NCBI translation strings are data (codon order       */
/* AAA, AAC, AAG, AAT, ACA, ...), transcribed from the NCBI genetic code definitions.    */
/* ------------------------------------------------------------------------------------ */
static const char *or_codes[34] = {
	0,
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF",
	"KNKNTTTT*S*SMIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSMIMIQHQHPPPPRRRRTTTTEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTSSSSMIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVQYQYSSSS*CWCLFLF",
	0, 0,
	"NNKNTTTTSSSSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSCCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLSLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF",
	"KNKNTTTTGSGSMIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"NNKNTTTTSSSSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVYY*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*YQYSSSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*YLYSSSS*CWCLFLF",
	0, 0, 0, 0,
	"NNKNTTTTSSSSMIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*YLY*SSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWC*FLF",
	"KNKNTTTTSSKSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSSGCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLALEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVQYQYSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVQYQYSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVYYYYSSSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVEYEYSSSS*CWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVEYEYSSSSWCWCLFLF",
	"KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*YWYSSSS*CWCLFLF",
	"KNKNTTTTSSKSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVVYY*YSSSSWCWCLFLF"
};
static const char or_nt_i2c[] = "ACGTN";
static const char or_aa_i2c[] = "ARNDCQEGHILKMFPSTWYV*X";

uint8_t or_nt4[256], or_aa20[256], or_codon[64];
/* BLOSUM62 over ARNDCQEGHILKMFPSTWYV*X (the published matrix; stop row/column -4, X per NCBI) */
int8_t or_blosum62[484] = {
	 4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-4, 0,
	-1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-4,-1,
	-2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3,-4,-1,
	-2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3,-4,-1,
	 0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-4,-2,
	-1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2,-4,-1,
	-1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2,-4,-1,
	 0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-4,-1,
	-2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3,-4,-1,
	-1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-4,-1,
	-1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-1,
	-1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2,-4,-1,
	-1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-4,-1,
	-2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-4,-1,
	-1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-4,-2,
	 1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2,-4, 0,
	 0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-4, 0,
	-3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-2,
	-2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-4,-1,
	 0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-4,-1,
	-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1,-4,
	 0,-1,-1,-1,-2,-1,-1,-1,-1,-1,-1,-1,-1,-1,-2, 0, 0,-2,-1,-1,-4,-1
};

/* nasw-tab.c:85-107 */
int or_make_tables(int codon_type)
{
	const char *tr;
	int i;
	if (codon_type < 0 || codon_type > 33) return -1;
	tr = or_codes[codon_type];
	if (!tr) return -2;
	memset(or_nt4, 4, 256);
	for (i = 0; or_nt_i2c[i]; ++i)
		or_nt4[i] = or_nt4[(uint8_t)toupper(or_nt_i2c[i])] = or_nt4[(uint8_t)tolower(or_nt_i2c[i])] = (uint8_t)i;
	memset(or_aa20, 21, 256);
	for (i = 0; or_aa_i2c[i]; ++i)
		or_aa20[i] = or_aa20[(uint8_t)toupper(or_aa_i2c[i])] = or_aa20[(uint8_t)tolower(or_aa_i2c[i])] = (uint8_t)i;
	for (i = 0; i < 64; ++i) or_codon[i] = or_aa20[(uint8_t)tr[i]];
	return 0;
}

/* nasw-tab.c:124-129 */
void or_opt_set_sp(ns_opt_t *o, int32_t model)
{
	static const int32_t mammal[6] = { 8, 15, 21, 30, 4, 4 }, generic[6] = { 8, 15, 21, 30, 0, 0 };
	int i;
	for (i = 0; i < 6; ++i)
		o->sp[i] = model == NS_S_MAMMAL? mammal[i] : model == NS_S_GENERIC? generic[i] : 0;
}

/* nasw-tab.c:131-147 */
void or_opt_init(ns_opt_t *o)
{
	memset(o, 0, sizeof(*o));
	o->go = 11, o->ge = 1, o->io = 29, o->fs = 17, o->xdrop = 100, o->end_bonus = 5;
	or_opt_set_sp(o, NS_S_MAMMAL);
	o->sp_null_bonus = -7, o->asize = 22, o->ie_coef = .5f;
	o->sc = or_blosum62, o->nt4 = or_nt4, o->aa20 = or_aa20, o->codon = or_codon;
}

/* nasw-tab.c:149-156: stop row/column set to -pen, stop/stop kept */
void or_set_stop_sc(int32_t asize, int8_t *mat, int8_t pen)
{
	int32_t s = or_aa20['*'], i, keep = mat[s * asize + s];
	for (i = 0; i < asize; ++i) mat[s * asize + i] = mat[i * asize + s] = (int8_t)-pen;
	mat[s * asize + s] = (int8_t)keep;
}

/* nasw-sse.c:341-349: polynomial log2 (exponent + quadratic of the mantissa), valid for x >= 2;
 * nasw-sse.c:758-761 and :446: penalty = (int)(ie_coef * log2(d) + .5f) for d >= 2. Evaluated with
 * separate float multiplies and adds (this file is built with -ffp-contract=off). */
static float or_log2(float x)
{
	union { float f; uint32_t i; } z;
	float e;
	z.f = x;
	e = (float)(int32_t)((z.i >> 23) & 255) - 128;
	z.i &= ~(255u << 23);
	z.i += 127u << 23;
	{
		float m = z.f, p = -0.34484843f * m;
		p = p + 2.02466578f;
		p = p * m;
		p = p - 0.67487759f;
		e = e + p;
	}
	return e;
}
int32_t or_ie_penalty(float ie_coef, int32_t d)
{
	float t;
	if (d < 2) return 0;
	t = ie_coef * or_log2((float)d);
	t = t + .5f;
	return (int32_t)t;
}

/* ------------------------------------------------------------------------------------ */
/* input preparation: nasw-sse.c:95-214                                                    */
/* ------------------------------------------------------------------------------------ */
static inline int8_t wrap8(int32_t v) { return (int8_t)(uint8_t)(v & 0xff); }

/* nasw-sse.c:95-108: nas[i] = residue of the codon ending at i; X unless 3 valid bases in a row */
static void or_translate(const char *ns, int32_t nl, const ns_opt_t *o, uint8_t *nas)
{
	int32_t i, run = 0;
	uint8_t c6 = 0;
	memset(nas, o->aa20['X'], (size_t)nl);
	for (i = 0; i < nl; ++i) {
		uint8_t c = o->nt4[(uint8_t)ns[i]];
		if (c < 4) {
			c6 = (uint8_t)((c6 << 2 | c) & 0x3f);
			if (++run >= 3) nas[i] = o->codon[c6];
		} else c6 = 0, run = 0;
	}
}

void or_prep(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int left, uint8_t *nas, uint8_t *aas, int8_t *donor, int8_t *acc)
{
	int32_t i, j;
	uint8_t *b = (uint8_t*)malloc((size_t)nl + 4); /* nt4 string as the kernel sees it */
	for (j = 0; j < al; ++j) aas[left? al - 1 - j : j] = o->aa20[(uint8_t)as[j]];
	for (i = 0; i < nl; ++i) b[left? nl - 1 - i : i] = o->nt4[(uint8_t)ns[i]];
	for (i = 0; i < nl + 1; ++i) donor[i] = acc[i] = wrap8(o->sp[3]);
	if (!left) { /* nasw-sse.c:125-150 */
		for (i = 0; i < nl - 3; ++i) {
			int32_t t = 3;
			if (b[i+1] == 2 && b[i+2] == 3) t = (i + 3 < nl && (b[i+3] == 0 || b[i+3] == 2))? (b[i] == 2? -1 : 4) : 0;
			else if (b[i+1] == 2 && b[i+2] == 1 && b[i] == 2) t = 1;
			else if (b[i+1] == 0 && b[i+2] == 3) t = 2;
			donor[i] = wrap8(t < 0? 0 : o->sp[t]);
		}
		for (i = 1; i < nl; ++i) {
			int32_t t = 3, pen_y = 0;
			if (b[i-1] == 0 && b[i] == 2) {
				t = (i >= 2 && (b[i-2] == 1 || b[i-2] == 3))? -1 : 0;
				for (j = i - 4; j >= 0 && j > i - 7; --j)
					if (b[j] != 1 && b[j] != 3) pen_y += o->sp[5];
			} else if (b[i-1] == 0 && b[i] == 1) t = 2;
			acc[i] = wrap8(t < 0? 0 : o->sp[t]);
			if (t == -1 || t == 0) acc[i] = wrap8(acc[i] + pen_y);
		}
		if (ss) { /* nasw-sse.c:142-157: ss[i] adjusts position i-1 */
			int32_t max_spsc = (o->io + 1) / 2 - 1;
			for (i = 1; i < nl; ++i) {
				int32_t sc = (int32_t)(int8_t)(ss[i] >> 1) - 64;
				if (sc > max_spsc) sc = max_spsc;
				if (ss[i] == 0xff) donor[i-1] = wrap8(donor[i-1] - o->sp_null_bonus), acc[i-1] = wrap8(acc[i-1] - o->sp_null_bonus);
				else if (ss[i] & 1) acc[i-1] = wrap8(acc[i-1] - sc);
				else donor[i-1] = wrap8(donor[i-1] - sc);
			}
		}
		or_translate(ns, nl, o, nas);
	} else { /* nasw-sse.c:170-213: reversed strings, donor/acceptor patterns mirrored */
		for (i = 0; i < nl - 3; ++i) {
			int32_t t = 3, pen_y = 0;
			if (b[i+1] == 2 && b[i+2] == 0) {
				t = (i + 3 < nl && (b[i+3] == 1 || b[i+3] == 3))? -1 : 0;
				for (j = i + 5; j < nl && j < i + 8; ++j)
					if (b[j] != 1 && b[j] != 3) pen_y += o->sp[5];
			} else if (b[i+1] == 1 && b[i+2] == 0) t = 2;
			donor[i] = wrap8(t < 0? 0 : o->sp[t]);
			if (t == -1 || t == 0) donor[i] = wrap8(donor[i] + pen_y);
		}
		for (i = 1; i < nl; ++i) {
			int32_t t = 3;
			if (b[i-1] == 3 && b[i] == 2) t = (i >= 2 && (b[i-2] == 0 || b[i-2] == 2))? (i + 1 < nl && b[i+1] == 2? -1 : 4) : 0;
			else if (b[i-1] == 1 && b[i] == 2 && i + 1 < nl && b[i+1] == 1) t = 1;
			else if (b[i-1] == 3 && b[i] == 0) t = 2;
			acc[i] = wrap8(t < 0? 0 : o->sp[t]);
		}
		if (ss) {
			int32_t max_spsc = (o->io + 1) / 2 - 1;
			for (i = 0; i < nl; ++i) {
				int32_t sc = (int32_t)(int8_t)(ss[i] >> 1) - 64, p = nl - i - 1;
				if (sc > max_spsc) sc = max_spsc;
				if (ss[i] == 0xff) donor[p] = wrap8(donor[p] - o->sp_null_bonus), acc[p] = wrap8(acc[p] - o->sp_null_bonus);
				else if (ss[i] & 1) donor[p] = wrap8(donor[p] - sc);
				else acc[p] = wrap8(acc[p] - sc);
			}
		}
		or_translate(ns, nl, o, b);   /* forward translation, then reverse and shift by two */
		for (i = 0; i < nl; ++i) nas[i] = b[nl - 1 - i];
		if (nl > 2) memmove(nas + 2, nas, (size_t)nl - 2);
		if (nl > 0) nas[0] = o->aa20['X'];
		if (nl > 1) nas[1] = o->aa20['X'];
	}
	free(b);
}

/* score of query column k (DP order) against residue a: nasw-sse.c:217-235 */
static inline int32_t or_score(const ns_opt_t *o, const uint8_t *aas, int32_t al, int32_t k, int32_t a)
{
	if (o->pssm) {
		int32_t q = (o->flag & NS_F_EXT_LEFT)? al - 1 - k : k;
		return o->pssm[q * o->asize + a];
	}
	return o->sc[a * o->asize + aas[k]];
}

/* ------------------------------------------------------------------------------------ */
/* 16-bit score-only DP, unsegmented column order (nasw-sse.c:19-26 recurrence; 360-459). */
/* Saturating int16 ops are monotone, so max commutes with them and the column order      */
/* gives the same H as the striped pass plus lazy-F.                                      */
/* ------------------------------------------------------------------------------------ */
static inline int32_t s16(int32_t x) { return x < -32768? -32768 : x > 32767? 32767 : x; }
#define ADDS(a, b) s16((a) + (b))
#define SUBS(a, b) s16((a) - (b))
#define MAX(a, b) ((a) > (b)? (a) : (b))

void or_dp16_score(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t npad, or_ext_t *out, int32_t *rows_done)
{
	const int32_t NEG = -32768, is_ext = !!(o->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	int32_t i, c, max_sc = INT32_MIN, max_sc_log = INT32_MIN, max_i = -1, pen_len = al * 3, last_i = 1;
	uint8_t *nas = (uint8_t*)malloc((size_t)nl + 1), *aas = (uint8_t*)malloc((size_t)al + 1);
	int8_t *donor = (int8_t*)malloc((size_t)nl + 2), *acc = (int8_t*)malloc((size_t)nl + 2);
	int32_t *Hr[4], *Dr[4], *A, *B, *C, *Hmax, *prof;
	for (c = 0; c < 4; ++c) {
		Hr[c] = (int32_t*)malloc(sizeof(int32_t) * (npad + 1));
		Dr[c] = (int32_t*)malloc(sizeof(int32_t) * npad);
		for (i = 0; i <= npad; ++i) Hr[c][i] = NEG;
		for (i = 0; i < npad; ++i) Dr[c][i] = NEG;
	}
	A = (int32_t*)malloc(sizeof(int32_t) * npad), B = (int32_t*)malloc(sizeof(int32_t) * npad), C = (int32_t*)malloc(sizeof(int32_t) * npad);
	Hmax = (int32_t*)malloc(sizeof(int32_t) * npad);
	prof = (int32_t*)malloc(sizeof(int32_t) * npad * o->asize);
	for (c = 0; c < npad; ++c) A[c] = B[c] = C[c] = Hmax[c] = NEG;
	or_prep(ns, nl, as, al, o, ss, !!(o->flag & NS_F_EXT_LEFT), nas, aas, donor, acc);
	{
		int32_t a;
		for (a = 0; a < o->asize; ++a)
			for (c = 0; c < npad; ++c) prof[a * npad + c] = c >= al? NEG : or_score(o, aas, al, c, a);
	}
	/* ring: Hr[(i)&3] holds row i (index 0 = column -1); rows -1, 0, 1 are neg_inf */
	for (i = 2; i < nl; ++i) {
		int32_t *H = Hr[i & 3] + 1, *H1 = Hr[(i - 1) & 3] + 1, *H2 = Hr[(i - 2) & 3] + 1, *H3 = Hr[(i - 3) & 3] + 1;
		int32_t *D = Dr[i & 3], *D3 = Dr[(i - 3) & 3];
		int32_t r = nas[i], gei = r == 20? o->fs : o->ge, *S = prof + r * npad;
		int32_t dm1 = donor[i-1], d0 = donor[i], dp1 = donor[i+1], a0 = acc[i], am1 = acc[i-1], am2 = acc[i-2];
		int32_t I = NEG, rowmax = NEG, tmp_sc, tmp_sc_log, end_sc;
		/* left boundary column -1: H(-1,-1)=0, H(0,-1)=H(1,-1)=-fs when computing row 2, else neg_inf */
		H3[-1] = i == 2? 0 : NEG;
		H2[-1] = i == 2? -o->fs : NEG;
		H1[-1] = i == 2? -o->fs : NEG;
		for (c = 0; c < npad; ++c) {
			int32_t h = ADDS(H3[c-1], S[c]), t, u;
			I = SUBS(MAX(SUBS(c > 0? H[c-1] : NEG, o->go), I), o->ge);
			h = MAX(h, I);
			t = SUBS(MAX(SUBS(H3[c], o->go), D3[c]), gei);
			D[c] = t, h = MAX(h, t);
			t = MAX(SUBS(SUBS(H1[c], o->io), dm1), A[c]);
			A[c] = t, h = MAX(h, SUBS(t, a0));
			h = MAX(h, SUBS(H1[c], o->fs));
			u = SUBS(H1[c-1], o->io);
			t = MAX(SUBS(u, d0), B[c]);
			B[c] = t, h = MAX(h, SUBS(t, am2));
			t = MAX(SUBS(u, dp1), C[c]);
			C[c] = t, h = MAX(h, SUBS(t, am1));
			h = MAX(h, SUBS(H1[c-1], o->fs));
			h = MAX(h, SUBS(H2[c], o->fs));
			h = MAX(h, SUBS(H2[c-1], o->fs));
			H[c] = h;
			rowmax = MAX(rowmax, h);
		}
		last_i = i;
		end_sc = H[al - 1] + o->end_bonus;
		tmp_sc = rowmax > end_sc? rowmax : end_sc;
		tmp_sc_log = tmp_sc - or_ie_penalty(o->ie_coef, i - pen_len);
		if (tmp_sc_log > max_sc_log) {
			max_sc = tmp_sc, max_sc_log = tmp_sc_log, max_i = i;
			memcpy(Hmax, H, sizeof(int32_t) * npad);
		}
		if (max_sc_log - tmp_sc_log > o->xdrop) break;
	}
	if (rows_done) *rows_done = last_i - 1;
	if (is_ext) { /* nasw-sse.c:448-458 */
		int32_t j, best_j = -1, best_sc = INT32_MIN;
		for (j = 0; j < al; ++j) {
			int32_t sc = Hmax[j];
			if (j == al - 1) sc += o->end_bonus;
			if (sc > best_sc) best_sc = sc, best_j = j;
			if (sc == max_sc) break;
		}
		if (j >= al) j = best_j, max_sc = best_sc;
		out->nt_len = max_i + 1, out->aa_len = j + 1, out->score = max_sc;
	} else {
		out->nt_len = nl, out->aa_len = al;
		out->score = nl > 2? Hr[last_i & 3][al] : NEG; /* H[al-1] of the last computed row (nasw-sse.c:559) */
	}
	for (c = 0; c < 4; ++c) free(Hr[c]), free(Dr[c]);
	free(A), free(B), free(C), free(Hmax), free(prof), free(nas), free(aas), free(donor), free(acc);
}

/* ------------------------------------------------------------------------------------ */
/* Striped global DP with trace, lane-by-lane emulation (nasw-sse.c:351-567, 779-973).   */
/* ------------------------------------------------------------------------------------ */
typedef struct { int32_t is32, vs, slen; int32_t neg; } or_sv_t;
static inline int32_t wrap32(int64_t x) { return (int32_t)(uint32_t)(uint64_t)x; }
static inline int32_t sv_add(const or_sv_t *s, int32_t a, int32_t b) { return s->is32? wrap32((int64_t)a + b) : s16(a + b); }
static inline int32_t sv_sub(const or_sv_t *s, int32_t a, int32_t b) { return s->is32? wrap32((int64_t)a - b) : s16(a - b); }

/* nasw-sse.c:34-42 */
static void or_fix_tiny_uv(int32_t n, uint32_t *cigar)
{
	int32_t i;
	for (i = 0; i < n; ++i) {
		int32_t op = cigar[i] & 0xf;
		if ((op == NS_CIGAR_U || op == NS_CIGAR_V) && cigar[i] >> 4 < 3) cigar[i] = cigar[i] >> 4 << 4 | NS_CIGAR_G;
	}
}

static uint32_t *or_push(uint32_t *cigar, int32_t *n, int32_t *m, uint32_t op, int32_t len)
{
	if (*n == 0 || op != (cigar[*n - 1] & 0xf) || op == NS_CIGAR_F || op == NS_CIGAR_G) {
		if (*n == *m) { *m += (*m >> 1) + 8; cigar = (uint32_t*)realloc(cigar, sizeof(uint32_t) * (size_t)*m); }
		cigar[(*n)++] = (uint32_t)len << 4 | op;
	} else cigar[*n - 1] += (uint32_t)len << 4;
	return cigar;
}

/* nasw-sse.c:44-93; tb[i][c] indexed by row i and column c */
static void or_backtrack(const int32_t *tb, int32_t npad, int32_t nl, int32_t al, ns_rst_t *r)
{
	int32_t i = nl - 1, j = al - 1, last = 0, k;
	uint32_t *cigar = r->cigar, tmp;
	while (i >= 2 && j >= 0) {
		int32_t x = tb[(size_t)i * npad + j], state, ext;
		if (x >> 9 & 1) x = 1 | x >> 4 << 4;
		state = last == 0? x & 0xf : last;
		ext = state >= 1 && state <= 5? x >> (state + 3) & 1 : 0;
		switch (state) {
		case 0: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_M, 1), i -= 3, --j; break;
		case 1: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_I, 1), --j; break;
		case 2: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_D, 1), i -= 3; break;
		case 3: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_N, 1), --i; break;
		case 4: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_U, 1), --i; if (!ext) --j; break;
		case 5: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_V, 1), --i; if (!ext) --j; break;
		case 6: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_F, 1), --i; break;
		case 7: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_F, 2), i -= 2; break;
		case 8: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_G, 1), --i, --j; break;
		case 9: cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_G, 2), i -= 2, --j; break;
		default: break; /* states 10..15 cannot occur */
		}
		last = state >= 1 && state <= 5 && ext? state : 0;
	}
	if (j > 0) cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_I, j);
	if (i >= 0) {
		int32_t l = (i + 1) / 3, t = (i + 1) % 3;
		if (l > 0) cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_D, l);
		if (t != 0) cigar = or_push(cigar, &r->n_cigar, &r->m_cigar, NS_CIGAR_F, t);
	}
	for (k = 0; k < r->n_cigar >> 1; ++k)
		tmp = cigar[k], cigar[k] = cigar[r->n_cigar - 1 - k], cigar[r->n_cigar - 1 - k] = tmp;
	or_fix_tiny_uv(r->n_cigar, cigar);
	r->cigar = cigar;
}

/* The striped CIGAR pass. Arrays of vectors are stored as v[j*vs + k] (stripe j, lane k);
 * column c = k*slen + j. Returns the trace in column order tbc[i*npad + c] and the score. */
static int32_t or_striped_cigar(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t is32, int32_t **tbc_out, int32_t *npad_out)
{
	or_sv_t s;
	int32_t vs, slen, npad, i, j, k, a, score;
	uint8_t *nas = (uint8_t*)malloc((size_t)nl + 1), *aas = (uint8_t*)malloc((size_t)al + 1);
	int8_t *donor = (int8_t*)malloc((size_t)nl + 2), *acc = (int8_t*)malloc((size_t)nl + 2);
	int32_t *ap, *Hb[4], *Db[4], *A, *B, *C, *tbc;
	s.is32 = is32, s.vs = vs = is32? 4 : 8, s.slen = slen = (al + vs - 1) / vs, s.neg = is32? -0x40000000 : -32768;
	npad = slen * vs;
	or_prep(ns, nl, as, al, o, ss, !!(o->flag & NS_F_EXT_LEFT), nas, aas, donor, acc);
	ap = (int32_t*)malloc(sizeof(int32_t) * (size_t)npad * o->asize);
	for (a = 0; a < o->asize; ++a) /* striped profile: nasw-sse.c:223-235 */
		for (j = 0; j < slen; ++j)
			for (k = 0; k < vs; ++k) {
				int32_t c = k * slen + j;
				ap[(size_t)a * npad + j * vs + k] = c >= al? s.neg : or_score(o, aas, al, c, a);
			}
	for (k = 0; k < 4; ++k) {
		Hb[k] = (int32_t*)malloc(sizeof(int32_t) * (size_t)(slen + 1) * vs); /* vector -1 first */
		Db[k] = (int32_t*)malloc(sizeof(int32_t) * (size_t)slen * vs);
		for (j = 0; j < (slen + 1) * vs; ++j) Hb[k][j] = s.neg;
		for (j = 0; j < slen * vs; ++j) Db[k][j] = s.neg;
	}
	A = (int32_t*)malloc(sizeof(int32_t) * (size_t)slen * vs), B = (int32_t*)malloc(sizeof(int32_t) * (size_t)slen * vs), C = (int32_t*)malloc(sizeof(int32_t) * (size_t)slen * vs);
	for (j = 0; j < slen * vs; ++j) A[j] = B[j] = C[j] = s.neg;
	tbc = (int32_t*)calloc((size_t)nl * npad, sizeof(int32_t));
	{
		/* H0..H3 as in NS_GEN_PREPARE; H3[-1][0] = 0, H2[-1][0] = H1[-1][0] = -fs (nasw-sse.c:262-264) */
		int32_t *H = Hb[0] + vs, *H1 = Hb[1] + vs, *H2 = Hb[2] + vs, *H3 = Hb[3] + vs;
		int32_t *D = Db[0], *D1 = Db[1], *D2 = Db[2], *D3 = Db[3], *tp;
		int32_t goe = o->go + o->ge;
		H3[-vs] = 0, H2[-vs] = -o->fs, H1[-vs] = -o->fs;
		for (i = 2; i < nl; ++i) {
			int32_t *S = ap + (size_t)nas[i] * npad, gei = nas[i] == 20? o->fs : o->ge;
			int32_t dim1 = donor[i-1], di = donor[i], dip1 = donor[i+1], ai = acc[i], aim1 = acc[i-1], aim2 = acc[i-2];
			int32_t I[8], last_h[8], z[8];
			for (k = 0; k < vs; ++k) I[k] = last_h[k] = s.neg;
			if (i > 2) { /* nasw-sse.c:276-282: shift the last stripe one lane up, neg_inf into lane 0 */
				for (k = vs - 1; k >= 1; --k) {
					H3[-vs + k] = H3[(slen - 1) * vs + k - 1];
					H2[-vs + k] = H2[(slen - 1) * vs + k - 1];
					H1[-vs + k] = H1[(slen - 1) * vs + k - 1];
				}
				H3[-vs] = H2[-vs] = H1[-vs] = s.neg;
			}
			for (j = 0; j < slen; ++j) {
				for (k = 0; k < vs; ++k) {
					int32_t h, t, u, v, y = 0, zz = 0, jk = j * vs + k, jm = (j - 1) * vs + k;
					h = sv_add(&s, H3[jm], S[jk]);
					t = sv_sub(&s, last_h[k], o->go);
					if (I[k] > t) zz |= 1 << 4;
					t = MAX(t, I[k]);
					I[k] = sv_sub(&s, t, o->ge);
					if (I[k] > h) y = 1;
					h = MAX(h, I[k]);
					u = sv_sub(&s, H3[jk], o->go), v = D3[jk];
					if (v > u) zz |= 1 << 5;
					t = sv_sub(&s, MAX(u, v), gei);
					D[jk] = t;
					if (t > h) y = 2;
					h = MAX(h, t);
					u = sv_sub(&s, H1[jk], o->io), v = A[jk];
					t = sv_sub(&s, u, dim1);
					if (v > t) zz |= 1 << 6;
					t = MAX(t, v), A[jk] = t, t = sv_sub(&s, t, ai);
					if (t > h) y = 3;
					h = MAX(h, t);
					u = sv_sub(&s, H1[jm], o->io), v = B[jk];
					t = sv_sub(&s, u, di);
					if (v > t) zz |= 1 << 7;
					t = MAX(t, v), B[jk] = t, t = sv_sub(&s, t, aim2);
					if (t > h) y = 4;
					h = MAX(h, t);
					v = C[jk];
					t = sv_sub(&s, u, dip1);
					if (v > t) zz |= 1 << 8;
					t = MAX(t, v), C[jk] = t, t = sv_sub(&s, t, aim1);
					if (t > h) y = 5;
					h = MAX(h, t);
					t = sv_sub(&s, H1[jk], o->fs); if (t > h) y = 6; h = MAX(h, t);
					t = sv_sub(&s, H2[jk], o->fs); if (t > h) y = 7; h = MAX(h, t);
					t = sv_sub(&s, H1[jm], o->fs); if (t > h) y = 8; h = MAX(h, t);
					t = sv_sub(&s, H2[jm], o->fs); if (t > h) y = 9; h = MAX(h, t);
					tbc[(size_t)i * npad + k * slen + j] = zz | y;
					H[jk] = h, last_h[k] = h;
				}
			}
			if (!is32) { /* lazy-F, int16: nasw-sse.c:537-553 */
				int32_t kk, brk = 0;
				for (k = 0; k < vs; ++k) I[k] = MAX(sv_sub(&s, last_h[k], goe), sv_sub(&s, I[k], o->ge));
				for (kk = 0; kk < vs && !brk; ++kk) {
					for (k = vs - 1; k >= 1; --k) I[k] = I[k-1];
					I[0] = s.neg;
					for (j = 0; j < slen; ++j) {
						int32_t all_le = 1;
						for (k = 0; k < vs; ++k) {
							int32_t jk = j * vs + k, h = H[jk];
							int32_t *zp = &tbc[(size_t)i * npad + k * slen + j];
							if (I[k] > h) *zp |= 1 << 9;
							h = MAX(h, I[k]);
							H[jk] = h;
							h = sv_sub(&s, h, goe);
							I[k] = sv_sub(&s, I[k], o->ge);
							if (I[k] > h) all_le = 0;
						}
						if (all_le) { brk = 1; break; }
					}
				}
			} else { /* lazy-F, int32: nasw-sse.c:937-960 */
				int32_t kk, brk = 0, itr[4];
				for (k = 0; k < vs; ++k) {
					int32_t h_open = sv_sub(&s, last_h[k], goe), i_ext = sv_sub(&s, I[k], o->ge);
					itr[k] = i_ext > h_open? 1 << 4 : 0;
					I[k] = MAX(h_open, i_ext);
				}
				for (kk = 0; kk < vs && !brk; ++kk) {
					for (k = vs - 1; k >= 1; --k) I[k] = I[k-1], itr[k] = itr[k-1];
					I[0] = s.neg, itr[0] = 0;
					for (j = 0; j < slen; ++j) {
						int32_t all_le = 1;
						for (k = 0; k < vs; ++k) {
							int32_t jk = j * vs + k, h = H[jk];
							int32_t *zp = &tbc[(size_t)i * npad + k * slen + j];
							if (I[k] > h) *zp = 1 | itr[k];
							h = MAX(h, I[k]);
							H[jk] = h;
							h = sv_sub(&s, h, goe);
							I[k] = sv_sub(&s, I[k], o->ge);
							itr[k] = 1 << 4;
							if (I[k] > h) all_le = 0;
						}
						if (all_le) { brk = 1; break; }
					}
				}
			}
			(void)z;
			tp = H3, H3 = H2, H2 = H1, H1 = H, H = tp;
			tp = D3, D3 = D2, D2 = D1, D1 = D, D = tp;
		}
		score = H1[((al - 1) % slen) * vs + (al - 1) / slen];
	}
	for (k = 0; k < 4; ++k) free(Hb[k]), free(Db[k]);
	free(A), free(B), free(C), free(ap), free(nas), free(aas), free(donor), free(acc);
	*tbc_out = tbc, *npad_out = npad;
	return score;
}

static void or_global_impl(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t is32, int32_t ext_width, ns_rst_t *r)
{
	int32_t is_ext = !!(o->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	r->n_cigar = 0, r->nt_len = nl, r->aa_len = al, r->score = INT32_MIN;
	if (!(o->flag & NS_F_CIGAR) || is_ext) { /* score-only paths */
		if (!is32) {
			int32_t vs = is_ext? ext_width : 8, npad = (al + vs - 1) / vs * vs;
			or_ext_t e;
			or_dp16_score(ns, nl, as, al, o, ss, npad, &e, 0);
			r->nt_len = e.nt_len, r->aa_len = e.aa_len, r->score = e.score;
		} else { /* gs32b score-only: no x-drop, H[al-1] of the last row */
			int32_t *tbc, npad;
			ns_opt_t o2 = *o;
			o2.flag = (o2.flag | NS_F_CIGAR) & ~(NS_F_EXT_LEFT | NS_F_EXT_RIGHT);
			if (o->flag & NS_F_EXT_LEFT) o2.flag |= NS_F_EXT_LEFT; /* keep the left preparation */
			r->score = or_striped_cigar(ns, nl, as, al, &o2, ss, 1, &tbc, &npad);
			free(tbc);
		}
		return;
	}
	{
		int32_t *tbc, npad;
		r->score = or_striped_cigar(ns, nl, as, al, o, ss, is32, &tbc, &npad);
		or_backtrack(tbc, npad, nl, al, r);
		free(tbc);
	}
}

void or_global16(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t ext_width, ns_rst_t *r)
{
	or_global_impl(ns, nl, as, al, o, ss, 0, ext_width, r);
}

void or_global32(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, ns_rst_t *r)
{
	or_global_impl(ns, nl, as, al, o, ss, 1, 8, r);
}

/* nasw-sse.c:980-999 without the GPU branch */
void or_global_auto(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t ext_width, ns_rst_t *r)
{
	int32_t i, max_sub = 0;
	int64_t max_score;
	if (o->pssm) { for (i = 0; i < al * o->asize; ++i) if (max_sub < o->pssm[i]) max_sub = o->pssm[i]; }
	else { for (i = 0; i < o->asize * o->asize; ++i) if (max_sub < o->sc[i]) max_sub = o->sc[i]; }
	max_score = (int64_t)al * max_sub + o->end_bonus;
	if (!(o->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) && max_score > 30000) or_global32(ns, nl, as, al, o, ss, r);
	else or_global16(ns, nl, as, al, o, ss, ext_width, r);
}

int32_t or_global_trace(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int32_t is32, uint16_t *trace_out, int32_t *score_out)
{
	int32_t *tbc, npad, k;
	*score_out = or_striped_cigar(ns, nl, as, al, o, ss, is32, &tbc, &npad);
	for (k = 0; k < nl * npad; ++k) trace_out[k] = (uint16_t)tbc[k];
	free(tbc);
	return npad;
}
