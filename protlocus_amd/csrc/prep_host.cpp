// Host-side input preparation for the ns_ext_prep() building block (nasw-sse.c:95-214):
// codon translation, donor/acceptor signal penalties and the optional splice-score track.
// The DP engines compute the same arrays on the device (ext_dp.hip: ext_prep_kernel).
#include <cstring>
#include <vector>
#include "nsh_internal.h"

static inline int8_t w8(int32_t v) { return (int8_t)(uint8_t)(v & 0xff); }

void nsh_prep_host(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *o,
		const uint8_t *ss, int left, uint8_t *nas, uint8_t *aas, int8_t *donor, int8_t *acc)
{
	std::vector<uint8_t> b(nl + 4);
	for (int32_t j = 0; j < al; ++j) aas[left? al - 1 - j : j] = o->aa20[(uint8_t)as[j]];
	for (int32_t i = 0; i < nl; ++i) b[left? nl - 1 - i : i] = o->nt4[(uint8_t)ns[i]];
	for (int32_t i = 0; i < nl + 1; ++i) donor[i] = acc[i] = w8(o->sp[3]);
	const int32_t max_spsc = (o->io + 1) / 2 - 1;
	auto spsc = [&](uint8_t s) { int32_t v = (int32_t)(int8_t)(s >> 1) - NS_SPSC_OFFSET; return v > max_spsc? max_spsc : v; };
	if (!left) {
		for (int32_t i = 0; i < nl - 3; ++i) {
			int32_t t = 3;
			if (b[i+1] == 2 && b[i+2] == 3) t = (i + 3 < nl && (b[i+3] == 0 || b[i+3] == 2))? (b[i] == 2? -1 : 4) : 0;
			else if (b[i+1] == 2 && b[i+2] == 1 && b[i] == 2) t = 1;
			else if (b[i+1] == 0 && b[i+2] == 3) t = 2;
			donor[i] = w8(t < 0? 0 : o->sp[t]);
		}
		for (int32_t i = 1; i < nl; ++i) {
			int32_t t = 3, pen_y = 0;
			if (b[i-1] == 0 && b[i] == 2) {
				t = (i >= 2 && (b[i-2] == 1 || b[i-2] == 3))? -1 : 0;
				for (int32_t j = i - 4; j >= 0 && j > i - 7; --j) if (b[j] != 1 && b[j] != 3) pen_y += o->sp[5];
			} else if (b[i-1] == 0 && b[i] == 1) t = 2;
			acc[i] = w8(t < 0? 0 : o->sp[t]);
			if (t == -1 || t == 0) acc[i] = w8(acc[i] + pen_y);
		}
		if (ss)
			for (int32_t i = 1; i < nl; ++i) {
				if (ss[i] == 0xff) donor[i-1] = w8(donor[i-1] - o->sp_null_bonus), acc[i-1] = w8(acc[i-1] - o->sp_null_bonus);
				else if (ss[i] & 1) acc[i-1] = w8(acc[i-1] - spsc(ss[i]));
				else donor[i-1] = w8(donor[i-1] - spsc(ss[i]));
			}
	} else {
		for (int32_t i = 0; i < nl - 3; ++i) {
			int32_t t = 3, pen_y = 0;
			if (b[i+1] == 2 && b[i+2] == 0) {
				t = (i + 3 < nl && (b[i+3] == 1 || b[i+3] == 3))? -1 : 0;
				for (int32_t j = i + 5; j < nl && j < i + 8; ++j) if (b[j] != 1 && b[j] != 3) pen_y += o->sp[5];
			} else if (b[i+1] == 1 && b[i+2] == 0) t = 2;
			donor[i] = w8(t < 0? 0 : o->sp[t]);
			if (t == -1 || t == 0) donor[i] = w8(donor[i] + pen_y);
		}
		for (int32_t i = 1; i < nl; ++i) {
			int32_t t = 3;
			if (b[i-1] == 3 && b[i] == 2) t = (i >= 2 && (b[i-2] == 0 || b[i-2] == 2))? (i + 1 < nl && b[i+1] == 2? -1 : 4) : 0;
			else if (b[i-1] == 1 && b[i] == 2 && i + 1 < nl && b[i+1] == 1) t = 1;
			else if (b[i-1] == 3 && b[i] == 0) t = 2;
			acc[i] = w8(t < 0? 0 : o->sp[t]);
		}
		if (ss)
			for (int32_t i = 0; i < nl; ++i) {
				const int32_t p = nl - i - 1;
				if (ss[i] == 0xff) donor[p] = w8(donor[p] - o->sp_null_bonus), acc[p] = w8(acc[p] - o->sp_null_bonus);
				else if (ss[i] & 1) donor[p] = w8(donor[p] - spsc(ss[i]));
				else acc[p] = w8(acc[p] - spsc(ss[i]));
			}
	}
	// residue of the codon ending at each DP position (left: translated forward, reversed, shifted by two)
	const uint8_t X = o->aa20[(uint8_t)'X'];
	std::vector<uint8_t> fwd(nl);
	for (int32_t i = 0, run = 0, c6 = 0; i < nl; ++i) {
		const uint8_t c = o->nt4[(uint8_t)ns[i]];
		fwd[i] = X;
		if (c < 4) { c6 = (c6 << 2 | c) & 0x3f; if (++run >= 3) fwd[i] = o->codon[c6]; }
		else c6 = 0, run = 0;
	}
	for (int32_t i = 0; i < nl; ++i) nas[i] = left? (i < 2? X : fwd[nl + 1 - i]) : fwd[i];
}
