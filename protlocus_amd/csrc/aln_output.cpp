// Alignment output: per-alignment statistics and CDS features (mp_extra_stop, mp_extra_start,
// mp_extra_cal; align.c:148-306) and the PAF / GFF3 / GTF records of mp_write_output
// (format.c:24-187, 333-477), from a CIGAR the DP engines returned. Host code: the DP engines
// hand back CIGARs, and these records are the last stage of the pipeline.
//
// Byte-identical to the reference's writers for the same alignment: the same integer
// accounting, the same "%.4f" formatting of identities, the same tag order. Two deliberate
// differences, neither visible in valid output: the CIGAR's span is checked before any base
// is read (the reference reads first and drops the alignment's details afterwards,
// align.c:258-268), and output goes to a caller buffer with snprintf semantics.
#include <cctype>
#include <cinttypes>
#include <cstdarg>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include "nasw-hip.h"
#include "nsh_internal.h"

namespace {

constexpr int kMpFGff = 0x8, kMpFNoPaf = 0x10, kMpFGtf = 0x20, kMpFShowResidue = 0x80, kMpFShowTrans = 0x100,
	kMpFNoCs = 0x200, kMpFAccuracy = 0x400;   // mp_mapopt_t.flag bits (protlocus.h:11-24)
constexpr int kMpAccRerank = 0x8;                  // mp_mapopt_t.acc_flags (protlocus.h:31)
constexpr int kFeatCds = 0, kFeatStop = 1;         // protlocus.h:43-44

// residue of the codon at nt[i..i+2]: X (aa20 'X') when any base is not ACGT
inline uint8_t codon_aa(const uint8_t *nt)
{
	return nt[0] > 3 || nt[1] > 3 || nt[2] > 3? ns_tab_aa20[(uint8_t)'X'] : ns_tab_codon[nt[0] << 4 | nt[1] << 2 | nt[2]];
}

// protein and target spans of a CIGAR, with the accounting of mp_extra_cal / mp_write_cs
void cigar_span(const uint32_t *cigar, int32_t n, int64_t *nl, int64_t *al)
{
	*nl = *al = 0;
	for (int32_t k = 0; k < n; ++k) {
		const int op = cigar[k] & 0xf;
		const int64_t len = cigar[k] >> 4;
		if (op == NS_CIGAR_M) *nl += 3 * len, *al += len;
		else if (op == NS_CIGAR_I) *al += len;
		else if (op == NS_CIGAR_D) *nl += 3 * len;
		else if (op == NS_CIGAR_F) *nl += len;
		else if (op == NS_CIGAR_G) *nl += len, ++*al;
		else if (op == NS_CIGAR_N) *nl += len;
		else if (op == NS_CIGAR_U || op == NS_CIGAR_V) *nl += len, ++*al;
	}
}

// distance from ve to the first in-frame stop codon inside [as, ae), -1 if none (align.c:283-293)
int32_t dist_to_stop(const nsh_aln_t *a, const uint8_t *nt, int64_t as, int64_t ae)
{
	for (int64_t j = a->ve; j + 2 < ae; j += 3)
		if (codon_aa(&nt[j - as]) == 20) return (int32_t)(j - a->ve);
	return -1;
}

// distance back from vs to an in-frame ATG before any stop codon, -1 if none (align.c:295-306)
int32_t dist_to_start(const nsh_aln_t *a, const uint8_t *nt, int64_t as, int64_t ae)
{
	for (int64_t j = a->vs; j >= as && j + 2 < ae; j -= 3) {
		const uint8_t aa = codon_aa(&nt[j - as]);
		if (aa == 20) break;
		if (aa == 12) return (int32_t)(a->vs - j);
	}
	return -1;
}

// appends printf-formatted text
void app(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void app(std::string &s, const char *fmt, ...)
{
	char buf[256];
	va_list ap;
	va_start(ap, fmt);
	const int l = vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	if (l < (int)sizeof(buf)) { s.append(buf, l); return; }
	std::string big(l + 1, '\0');
	va_start(ap, fmt);
	vsnprintf(&big[0], l + 1, fmt, ap);
	va_end(ap);
	s.append(big.data(), l);
}

inline char nt_lc(uint8_t c) { return "acgtn"[c > 4? 4 : c]; }

// the cs:Z: tag body (format.c:102-187); nt holds the bases of [vs, ve), aa the protein from qs
void write_cs(std::string &s, const nsh_aln_t *a, const uint8_t *nt, const char *aa)
{
	int64_t nl = 0, al = 0;
	s += "cs:Z:";
	for (int32_t k = 0; k < a->n_cigar; ++k) {
		const int op = a->cigar[k] & 0xf;
		const int64_t len = a->cigar[k] >> 4, len3 = 3 * len;
		if (op == NS_CIGAR_M) {
			int64_t t = 0;
			for (int64_t l = 0, i = nl, j = al; l < len; ++l, ++j, i += 3) {
				const uint8_t nt_aa = codon_aa(&nt[i]), aa_aa = ns_tab_aa20[(uint8_t)aa[j]];
				if (nt_aa != aa_aa) {
					if (t > 0) app(s, ":%" PRId64, t);
					s += '*', s += nt_lc(nt[i]), s += nt_lc(nt[i + 1]), s += nt_lc(nt[i + 2]);
					s += (char)toupper((unsigned char)aa[j]);
					t = 0;
				} else ++t;
			}
			if (t > 0) app(s, ":%" PRId64, t);
			nl += len3, al += len;
		} else if (op == NS_CIGAR_I) {
			s += '+';
			for (int64_t j = 0; j < len; ++j) s += (char)toupper((unsigned char)aa[al + j]);
			al += len;
		} else if (op == NS_CIGAR_D) {
			s += '-';
			for (int64_t i = 0; i < len3; ++i) s += nt_lc(nt[nl + i]);
			nl += len3;
		} else if (op == NS_CIGAR_F) {
			s += '-';
			for (int64_t i = 0; i < len; ++i) s += nt_lc(nt[nl + i]);
			nl += len;
		} else if (op == NS_CIGAR_G) {
			s += '*';
			for (int64_t i = 0; i < len; ++i) s += nt_lc(nt[nl + i]);
			s += (char)toupper((unsigned char)aa[al]);
			nl += len, ++al;
		} else if (op == NS_CIGAR_N || op == NS_CIGAR_U || op == NS_CIGAR_V) {
			const int64_t lsh = op == NS_CIGAR_N? 0 : op == NS_CIGAR_U? 1 : 2, rsh = lsh == 0? 0 : 3 - lsh;
			if (lsh > 0) {   // the split codon's first bases and its residue
				s += '*';
				for (int64_t i = 0; i < lsh; ++i) s += nt_lc(nt[nl + i]);
				s += (char)toupper((unsigned char)aa[al]);
			}
			app(s, "~%c%c%" PRId64 "%c%c", nt_lc(nt[nl + lsh]), nt_lc(nt[nl + lsh + 1]), len - (lsh + rsh),
					nt_lc(nt[nl + len - rsh - 2]), nt_lc(nt[nl + len - rsh - 1]));
			if (rsh > 0) {
				s += '-';
				for (int64_t i = 0; i < rsh; ++i) s += nt_lc(nt[nl + len - rsh + i]);
			}
			if (lsh) ++al;
			nl += len;
		}
	}
}

std::string fmt4(double x)
{
	char dec[16];
	snprintf(dec, sizeof(dec), "%.4f", x);
	return dec;
}

} // namespace

extern "C" int32_t nsh_aln_stats(const nsh_aln_t *a, const uint8_t *nt, int64_t as, int64_t ae, const ns_opt_t *opt,
		const int8_t *pssm, nsh_aln_stats_t *st, nsh_feat_t *feat, int32_t m_feat)
{
	if (!a || !nt || !opt || !st || !a->aa || (a->n_cigar > 0 && !a->cigar) || a->n_cigar < 0 || !opt->sc ||
			a->qs < 0 || a->qe < a->qs || a->qe > a->qlen || as > a->vs || a->vs > a->ve || a->ve > ae) {
		nsh_set_error("nsh_aln_stats: bad arguments");
		return -1;
	}
	const int32_t asize = opt->asize;
	memset(st, 0, sizeof(*st));
	st->dist_stop = dist_to_stop(a, nt, as, ae);
	st->dist_start = dist_to_start(a, nt, as, ae);
	int64_t span_nl, span_al;
	cigar_span(a->cigar, a->n_cigar, &span_nl, &span_al);
	if (span_nl != a->ve - a->vs || span_al != a->qe - a->qs) return 0;   // details dropped (align.c:258-268)
	int32_t n_intron = 0;
	for (int32_t k = 0; k < a->n_cigar; ++k) {
		const int op = a->cigar[k] & 0xf;
		n_intron += op == NS_CIGAR_N || op == NS_CIGAR_U || op == NS_CIGAR_V;
	}
	const bool has_stop = a->qe == a->qlen && st->dist_stop == 0;
	st->n_exon = n_intron + 1;
	st->n_feat = st->n_exon + has_stop;
	if (!feat || m_feat < st->n_feat) {
		nsh_set_error("nsh_aln_stats: %d features do not fit in %d", st->n_feat, m_feat);
		return -1;
	}
	memset(feat, 0, sizeof(nsh_feat_t) * st->n_feat);
	const uint8_t *v = nt + (a->vs - as);        // bases from vs
	const int64_t l_v = ae - a->vs;
	const char *q = a->aa + a->qs;               // protein from qs
	const uint8_t aa_ambi = ns_tab_aa20[(uint8_t)'X'], aa_stop = ns_tab_aa20[(uint8_t)'*'];
	auto score = [&](int64_t j, uint8_t nt_aa, uint8_t aa_aa) -> int32_t {
		return pssm? pssm[j * asize + nt_aa] : opt->sc[nt_aa * asize + aa_aa];
	};
	int64_t nl = 0, al = 0, vs0 = a->vs;
	int32_t blen0 = 0, n_iden0 = 0, score0 = 0, n_fs0 = 0, n_stop0 = 0, phase0 = 0, qs0 = a->qs, ft = 0;
	char acc0[2] = { 0, 0 };
	for (int32_t k = 0; k < a->n_cigar; ++k) {
		const int op = a->cigar[k] & 0xf;
		const int64_t len = a->cigar[k] >> 4;
		if (op == NS_CIGAR_M) {
			for (int64_t l = 0, i = nl, j = al; l < len; ++l, ++j, i += 3) {
				const uint8_t nt_aa = codon_aa(&v[i]), aa_aa = ns_tab_aa20[(uint8_t)q[j]];
				const int32_t s = score(j, nt_aa, aa_aa);
				st->n_stop += nt_aa == aa_stop, st->n_iden += nt_aa == aa_aa, st->n_plus += s > 0, st->dp_max += s;
			}
			nl += 3 * len, al += len, st->blen += (int32_t)(3 * len);
		} else if (op == NS_CIGAR_I) {
			st->dp_max -= opt->go + opt->ge * (int32_t)len;
			al += len, st->blen += (int32_t)(3 * len);
		} else if (op == NS_CIGAR_D) {
			for (int64_t l = 0, i = nl; l < len; ++l, i += 3)   // in-frame stops inside deletions count
				st->n_stop += codon_aa(&v[i]) == aa_stop;
			st->dp_max -= opt->go + opt->ge * (int32_t)len;
			nl += 3 * len, st->blen += (int32_t)(3 * len);
		} else if (op == NS_CIGAR_F) {
			st->dp_max -= opt->fs;
			nl += len, st->blen += (int32_t)len, ++st->n_fs;
		} else if (op == NS_CIGAR_G) {
			st->dp_max -= opt->fs;
			nl += len, ++al, st->blen += 3, ++st->n_fs;
		} else if (op == NS_CIGAR_N || op == NS_CIGAR_U || op == NS_CIGAR_V) {
			if (op != NS_CIGAR_N) {   // the codon split by a phase-1 or phase-2 intron
				const uint8_t n1 = v[nl], n2 = op == NS_CIGAR_U? v[nl + len - 2] : v[nl + 1], n3 = v[nl + len - 1];
				const uint8_t nt_aa = n1 > 3 || n2 > 3 || n3 > 3? aa_ambi : ns_tab_codon[n1 << 4 | n2 << 2 | n3];
				const uint8_t aa_aa = ns_tab_aa20[(uint8_t)q[al]];
				const int32_t s = score(al, nt_aa, aa_aa);
				st->n_stop += nt_aa == aa_stop, st->n_iden += nt_aa == aa_aa, st->n_plus += s > 0, st->dp_max += s;
				st->blen += 3;
			}
			nsh_feat_t &f = feat[ft++];   // the exon this intron closes
			f.type = kFeatCds, f.vs = vs0, f.qs = qs0, f.qe = (int32_t)(a->qs + al), f.phase = (int16_t)phase0;
			f.blen = st->blen - blen0, f.n_iden = st->n_iden - n_iden0, f.n_fs = st->n_fs - n_fs0;
			f.n_stop = st->n_stop - n_stop0, f.score = st->dp_max - score0;
			if (ft > 1) f.acceptor[0] = acc0[0], f.acceptor[1] = acc0[1];
			if (op == NS_CIGAR_N) f.ve = a->vs + nl, vs0 = a->vs + nl + len, phase0 = 0;
			else if (op == NS_CIGAR_U) f.ve = a->vs + nl + 1, vs0 = a->vs + nl + len - 2, phase0 = 2;
			else f.ve = a->vs + nl + 2, vs0 = a->vs + nl + len - 1, phase0 = 1;
			f.donor[0] = f.ve - a->vs < l_v? ns_tab_nt_i2c[v[f.ve - a->vs]] : '.';
			f.donor[1] = f.ve - a->vs + 1 < l_v? ns_tab_nt_i2c[v[f.ve - a->vs + 1]] : '.';
			qs0 = f.qe, n_fs0 = st->n_fs, n_stop0 = st->n_stop, score0 = st->dp_max, blen0 = st->blen, n_iden0 = st->n_iden;
			acc0[0] = vs0 - a->vs >= 2? ns_tab_nt_i2c[v[vs0 - a->vs - 2]] : '.';
			acc0[1] = vs0 - a->vs >= 1? ns_tab_nt_i2c[v[vs0 - a->vs - 1]] : '.';
			nl += len, al += op != NS_CIGAR_N;
		}
	}
	nsh_feat_t &f = feat[ft++];   // the last exon
	f.type = kFeatCds, f.vs = vs0, f.ve = a->vs + nl, f.qs = qs0, f.qe = (int32_t)(a->qs + al), f.phase = (int16_t)phase0;
	f.blen = st->blen - blen0, f.n_iden = st->n_iden - n_iden0, f.n_fs = st->n_fs - n_fs0;
	f.n_stop = st->n_stop - n_stop0, f.score = st->dp_max - score0;
	if (ft > 1) f.acceptor[0] = acc0[0], f.acceptor[1] = acc0[1];
	if (has_stop) {   // the stop codon after ve (its protein coordinates as the reference sets them)
		nsh_feat_t &g = feat[ft++];
		g.type = kFeatStop, g.vs = a->ve, g.ve = a->ve + 3, g.qs = g.qe = (int32_t)(a->qe + al);
		g.phase = 0, g.n_fs = 0, g.blen = 3, g.n_iden = 0;
	}
	return st->n_feat;
}

extern "C" int64_t nsh_write_output(char *out, int64_t size, const nsh_aln_t *a, const nsh_aln_stats_t *st,
		const nsh_feat_t *feat, const uint8_t *nt, int64_t as, int32_t flag, int32_t acc_flags, const char *gff_prefix,
		int64_t id, int32_t gff_delim, int32_t hit_idx)
{
	if (!a || !a->qname || !a->ctg_name || !a->aa || size < 0 || (size > 0 && !out) || (st && (!feat || !nt)) ||
			(st && a->vs < as)) {
		nsh_set_error("nsh_write_output: bad arguments");
		return -1;
	}
	if (flag & (kMpFShowResidue | kMpFShowTrans)) {
		nsh_set_error("nsh_write_output: residue/translation display (MP_F_SHOW_RESIDUE, MP_F_SHOW_TRANS) is not supported");
		return -1;
	}
	if (!gff_prefix) gff_prefix = "PL";
	std::string s;
	const char strand = "+-"[a->rev & 1];
	const int64_t L = a->ctg_len;
	const bool has_stop = st && a->qe == a->qlen && st->dist_stop == 0;
	const int64_t ve_mrna = has_stop? a->ve + 3 : a->ve;
	auto paf = [&]() {   // format.c:333-362
		if (flag & (kMpFGff | kMpFGtf)) s += "##PAF\t";
		app(s, "%s\t%d\t%d\t%d\t%c\t%s\t%ld\t", a->qname, a->qlen, a->qs, a->qe, strand, a->ctg_name, (long)L);
		if (a->rev & 1) app(s, "%ld\t%ld\t", (long)(L - a->ve), (long)(L - a->vs));
		else app(s, "%ld\t%ld\t", (long)a->vs, (long)a->ve);
		if (st) {
			app(s, "%d\t%d\t0\tAS:i:%d\tms:i:%d\tnp:i:%d\tfs:i:%d\tst:i:%d\tda:i:%d\tdo:i:%d\t", st->n_iden * 3, st->blen,
					a->dp_score, st->dp_max, st->n_plus, st->n_fs, st->n_stop, st->dist_start, st->dist_stop);
			if (a->profile_score != INT32_MIN) app(s, "ps:i:%d\t", a->profile_score);
			if ((flag & kMpFAccuracy) && (acc_flags & kMpAccRerank)) app(s, "ra:i:%d\t", a->rank_sc);
			s += "cg:Z:";
			for (int32_t k = 0; k < a->n_cigar; ++k) app(s, "%u%c", a->cigar[k] >> 4, NS_CIGAR_STR[a->cigar[k] & 0xf]);
		} else app(s, "%d\t%d\t%d", a->chn_sc, a->chn_sc_ungap, a->cnt);
		if (!(flag & kMpFNoCs)) {
			s += '\t';
			if (st) write_cs(s, a, nt + (a->vs - as), a->aa + a->qs);
		}
		s += '\n';
	};
	auto gff = [&]() {   // format.c:364-416
		if (!st) return;
		std::string ids;
		if (gff_delim >= 33 && gff_delim <= 126 && hit_idx >= 0) app(ids, "%s%c%d", a->qname, gff_delim, hit_idx);
		else app(ids, "%s%.6ld", gff_prefix, (long)id);
		const int64_t vs = a->rev & 1? L - ve_mrna : a->vs, ve = a->rev & 1? L - a->vs : ve_mrna;
		app(s, "%s\tProtLocus\tmRNA\t%ld\t%ld\t%d\t%c\t.\tID=%s;Rank=%d", a->ctg_name, (long)vs + 1, (long)ve, st->dp_max,
				strand, ids.c_str(), hit_idx);
		s += ";Identity=" + fmt4((double)st->n_iden * 3 / st->blen);
		s += ";Positive=" + fmt4((double)st->n_plus * 3 / st->blen);
		if (st->n_fs > 0) app(s, ";Frameshift=%d", st->n_fs);
		if (st->n_stop > 0) app(s, ";StopCodon=%d", st->n_stop);
		app(s, ";Target=%s %d %d\n", a->qname, a->qs + 1, a->qe);
		for (int32_t j = 0; j < st->n_feat; ++j) {
			const nsh_feat_t &f = feat[j];
			int64_t fve = f.ve;
			if (has_stop && f.type == kFeatCds && j + 1 < st->n_feat && feat[j + 1].type == kFeatStop) fve += 3;   // the last CDS includes the stop codon
			const int64_t gvs = a->rev & 1? L - fve : f.vs, gve = a->rev & 1? L - f.vs : fve;
			app(s, "%s\tProtLocus\t%s\t%ld\t%ld\t%d\t%c\t%d\tParent=%s;Rank=%d", a->ctg_name,
					f.type == kFeatStop? "stop_codon" : "CDS", (long)gvs + 1, (long)gve, f.score, strand, f.phase, ids.c_str(), hit_idx);
			if (f.type == kFeatCds) {
				s += ";Identity=" + fmt4((double)f.n_iden * 3 / f.blen);
				if (f.acceptor[0] && strncmp(f.acceptor, "AG", 2) != 0) app(s, ";Acceptor=%c%c", f.acceptor[0], f.acceptor[1]);
				if (f.donor[0] && strncmp(f.donor, "GT", 2) != 0) app(s, ";Donor=%c%c", f.donor[0], f.donor[1]);
				if (f.n_fs > 0) app(s, ";Frameshift=%d", f.n_fs);
				if (f.n_stop > 0) app(s, ";StopCodon=%d", f.n_stop);
				app(s, ";Target=%s %d %d", a->qname, f.qs + 1, f.qe);
			}
			s += '\n';
		}
	};
	auto gtf = [&]() {   // format.c:418-455
		if (!st) return;
		std::string gid, tid;
		app(gid, "%sG%.6ld", gff_prefix, (long)id), app(tid, "%sT%.6ld", gff_prefix, (long)id);
		const int64_t vs = a->rev & 1? L - ve_mrna : a->vs, ve = a->rev & 1? L - a->vs : ve_mrna;
		app(s, "%s\tProtLocus\tgene\t%ld\t%ld\t%d\t%c\t.\tgene_id \"%s\";\n", a->ctg_name, (long)vs + 1, (long)ve, st->dp_max,
				strand, gid.c_str());
		app(s, "%s\tProtLocus\ttranscript\t%ld\t%ld\t%d\t%c\t.\ttranscript_id \"%s\"; gene_id \"%s\";\n", a->ctg_name,
				(long)vs + 1, (long)ve, st->dp_max, strand, tid.c_str(), gid.c_str());
		for (int32_t j = 0; j < st->n_feat; ++j) {
			const nsh_feat_t &f = feat[j];
			if (f.type != kFeatCds) continue;
			int64_t cvs = a->rev & 1? L - f.ve : f.vs, cve = a->rev & 1? L - f.vs : f.ve, evs = cvs, eve = cve;
			if (f.ve == a->ve) {   // the last exon reaches the stop codon
				if (a->rev & 1) evs = L - ve_mrna;
				else eve = ve_mrna;
			}
			app(s, "%s\tProtLocus\texon\t%ld\t%ld\t%d\t%c\t.\ttranscript_id \"%s\"; gene_id \"%s\";\n", a->ctg_name,
					(long)evs + 1, (long)eve, f.score, strand, tid.c_str(), gid.c_str());
			app(s, "%s\tProtLocus\tCDS\t%ld\t%ld\t%d\t%c\t%d\ttranscript_id \"%s\"; gene_id \"%s\";\n", a->ctg_name,
					(long)cvs + 1, (long)cve, f.score, strand, f.phase, tid.c_str(), gid.c_str());
		}
	};
	if (flag & kMpFGtf) gtf();   // format.c:457-477
	else {
		if (!(flag & kMpFNoPaf)) paf();
		if (flag & kMpFGff) gff();
	}
	if (size > 0) {
		const int64_t l = (int64_t)s.size() < size - 1? (int64_t)s.size() : size - 1;
		memcpy(out, s.data(), l);
		out[l] = 0;
	}
	return (int64_t)s.size();
}
