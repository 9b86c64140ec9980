// Device side of the batched extension DP (execution model: see ext_dp.hip).
// Included by the ext_dp_k*.hip translation units, one per column-count group.
#ifndef NSH_EXT_DP_IMPL_H
#define NSH_EXT_DP_IMPL_H
#include <hip/hip_runtime.h>
#include <climits>
#include <type_traits>
#include "nsh_internal.h"
#include "ext_dp_types.h"

namespace nsh_ext {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
#define DEV static __device__ __forceinline__

// saturating int16 pairs: v_pk_add_i16/v_pk_sub_i16 with clamp, v_pk_max_i16, v_pk_min_i16
DEV s2 adds(s2 a, s2 b) { return __builtin_elementwise_add_sat(a, b); }
DEV s2 subs(s2 a, s2 b) { return __builtin_elementwise_sub_sat(a, b); }
DEV s2 vmax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
DEV s2 vmin(s2 a, s2 b) { return __builtin_elementwise_min(a, b); }
// the same values biased by +32768 as unsigned pairs, where 0 stands for -32768 (scan only)
DEV u2 subu(u2 a, u2 b) { return __builtin_elementwise_sub_sat(a, b); }
DEV u2 maxu(u2 a, u2 b) { return __builtin_elementwise_max(a, b); }
DEV int pk(s2 a) { return __builtin_bit_cast(int, a); }
DEV s2 up(int a) { return __builtin_bit_cast(s2, a); }
DEV int pku(u2 a) { return __builtin_bit_cast(int, a); }
DEV u2 upu(int a) { return __builtin_bit_cast(u2, a); }
DEV s2 spl(int x) { return s2{(short)x, (short)x}; }
DEV u2 splu(int x) { return u2{(unsigned short)x, (unsigned short)x}; }
DEV int lo16(int x) { return (int)(short)(x & 0xffff); }
DEV int hi16(int x) { return (int)(short)((unsigned)x >> 16); }

constexpr int NEG = -32768;
constexpr int PACKNEG = (int)0x80008000u;
constexpr int BIAS = (int)0x80008000u;   // xor: signed pair <-> biased unsigned pair
constexpr int NRES = 22;                 // residue codes of the profile rows

// DPP controls (GFX9 encoding)
enum : int { ROW_SHR1 = 0x111, ROW_SHR2 = 0x112, ROW_SHR4 = 0x114, ROW_SHR8 = 0x118,
	WAVE_SHR1 = 0x138, WAVE_ROR1 = 0x13C, ROW_BCAST15 = 0x142, ROW_BCAST31 = 0x143 };
// lanes without a source read 0 (bound_ctrl); every lane is written
template <int CTRL>
DEV int dpp0(int x) { return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true); }
// lanes without a source keep their own value
template <int CTRL>
DEV int dppk(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false); }

// PROF: scores are staged per problem as an int8 profile (22 residues x 128C columns) in LDS,
// from the PSSM rows or the matrix; otherwise (matrix only) the 22 x 22 matrix is staged once per
// work-group as an int16 table of 22 row residues x 23 column residues (column 22 = padding,
// -32768), which leaves the LDS small enough for register-bound occupancy. Per row, lane j < 23
// reads the entry (row residue, j) into one register, and every column takes its score from the
// lane of its residue with ds_bpermute: the lanes' source addresses are fixed for the problem, so
// a column's score costs no address arithmetic, only the perm that packs the two halves.
constexpr int TAB_PAD = 22;    // column residue of the padded columns (score -32768)
constexpr int TAB_W = 23;      // entries per table row

// uniform loads through the scalar data cache (the row records and the penalty table are not
// written while the kernels run)
typedef const uint32_t __attribute__((address_space(4))) *cu32_t;
typedef const int32_t __attribute__((address_space(4))) *ci32_t;
DEV int sld(const uint32_t *p, int64_t i) { return (int)((cu32_t)p)[i]; }
DEV int sld(const int32_t *p, int64_t i) { return ((ci32_t)p)[i]; }
typedef const uint16_t __attribute__((address_space(3))) *lds_u16_t;   // 32-bit LDS addressing

template <int C, bool PROF>
DEV void ext_solve(const ExtArgs &a, const ExtProb &P, int8_t *prof, int *snap, int *dring, const int L)
{
	constexpr int HALF = 64 * C, ROWB = 128 * C;
	// npad is a multiple of 8 (the host checks it), so when C divides 8 a lane's columns of one half
	// are either all below npad or all at or beyond it: the row maximum is masked once per lane
	constexpr bool LANE_MASK = (8 % C) == 0;
	const int nl = P.nl, al = P.al, npad = P.npad, asz = a.asize;
	const int go = a.go, ge = a.ge, fs = a.fs, io = P.io, eb = a.eb, xdrop = a.xdrop;
	const uint8_t *aas = a.aas + P.aa_off;
	const int8_t *pssm = P.pssm_off >= 0? a.pssm + P.pssm_off : nullptr;
	const uint32_t *rec = a.rec + P.rec_off;
	const int pen_len = 3 * al;

	// ---- per-lane column constants and the scores ----
	// sel[k] (profile path): v_perm selector packing the two sign-extended profile bytes of column k into an
	// int16 pair; padded columns (al <= c < npad) become -32768 like the reference profile.
	// lim[k]: columns at or beyond npad (not part of the reference's row maximum) are masked.
	uint32_t sel[PROF? C : 1];
	s2 lim[LANE_MASK? 1 : C];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		const int clo = L * C + k, chi = HALF + L * C + k;
		const bool plo = clo >= al && clo < npad, phi = chi >= al && chi < npad;
		if (PROF) sel[k] = (plo? 0x000Cu : 0x0100u) | (phi? 0x040C0000u : 0x05040000u);
		if (!LANE_MASK) lim[k] = s2{(short)(clo < npad? 32767 : NEG), (short)(chi < npad? 32767 : NEG)};
	}
	if (LANE_MASK) lim[0] = s2{(short)(L * C < npad? 32767 : NEG), (short)(HALF + L * C < npad? 32767 : NEG)};
	// table path: ds_bpermute source addresses (column residue * 4, padding past al) of the lane's
	// columns, low and high half
	int bl[PROF? 1 : C], bh[PROF? 1 : C];
	if (PROF) {
		__syncthreads();
#pragma unroll
		for (int h = 0; h < 2; ++h)
#pragma unroll
			for (int k = 0; k < C; ++k) {
				const int c = h * HALF + L * C + k;
				const bool real = c < al;
				const int q = P.left? al - 1 - c : c;
				int8_t *dst = prof + h * HALF + k * 64 + L;
				const int aa = real? aas[c] : 0;
				for (int r = 0; r < NRES; ++r)
					dst[r * ROWB] = (int8_t)(real? (pssm? pssm[q * asz + r] : a.mat[r * asz + aa]) : -128);
			}
		__syncthreads();
	} else {
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int clo = L * C + k, chi = HALF + L * C + k;
			bl[k] = (clo < al? aas[clo] : TAB_PAD) * 4;
			bh[k] = (chi < al? aas[chi] : TAB_PAD) * 4;
		}
	}
	const int trow_col = L < TAB_W? L : TAB_PAD;   // the lane's entry in a table row
	const lds_u16_t tab = (lds_u16_t)prof;

	// ---- scan constants (biased unsigned): the insertion carry decays by ge per column ----
	const int geC = ge * C;
	const u2 DEC1 = splu(geC), DEC2 = splu(2 * geC), DEC4 = splu(4 * geC), DEC8 = splu(8 * geC);
	const u2 DEC15 = splu(geC * ((L & 15) + 1));                  // row_bcast:15 source distance
	const u2 DEC31 = splu(L >= 32? geC * (L - 31) : 0xffff);       // row_bcast:31 (rows 0, 1: no source)
	const u2 DECL = u2{0, (unsigned short)(geC * L)};              // low half's total into the high half
	const s2 GO = spl(go), GE = spl(ge), IO = spl(io), FS = spl(fs);
	// v_perm selector of the left-neighbour exchange: lane 0 takes -32768 in the low half and the
	// rotated-in value's low half (lane 63's last low column) in the high half
	const uint32_t nsel = L == 0? 0x05040100u : 0x07060504u;

	// ---- DP state: three rows of H and D, intron states, best-row snapshot ----
	// (the best row's H values are kept in LDS: snap[k * 64 + L], written on each new best row)
	// (D of rows i-1..i-3 lives in LDS: dring[(row % 3) * C + k][L])
	s2 Hs[3][C], A[C], B[C], Cc[C], nh[3];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		Hs[0][k] = Hs[1][k] = Hs[2][k] = up(PACKNEG);
		dring[(0 * C + k) * 64 + L] = dring[(1 * C + k) * 64 + L] = dring[(2 * C + k) * 64 + L] = PACKNEG;
		A[k] = B[k] = Cc[k] = up(PACKNEG);
		snap[k * 64 + L] = PACKNEG;
	}
	// left neighbour column of each lane's first column, rows i-3 (nh[0]), i-2 (nh[1]), i-1 (nh[2])
	// at the first row; column -1 holds H(-1,-1)=0 and H(0,-1)=H(1,-1)=-fs only for row 2.
	nh[0] = nh[1] = nh[2] = up(PACKNEG);
	if (L == 0) nh[0] = s2{0, (short)NEG}, nh[1] = nh[2] = s2{(short)-fs, (short)NEG};

	// end column (al-1): half, lane and slot
	const int he = (al - 1) >= HALF, qe = (al - 1) - he * HALF, Le = qe / C, ke = qe % C;

	// ---- row records and penalties: scalar loads issued a row before they are needed ----
	auto pen_of = [&](int r) -> int {
		const int d = r - pen_len;
		return d < 2? 0 : sld(a.pen, d < a.pen_n? d : a.pen_n - 1);
	};
	int d0 = 0, dp1 = 0, am1 = 0, a0 = 0, dm1 = 0, am2 = 0;
	int rc1 = 0, rc2 = 0;          // records of rows i and i+1 on entry to row i
	int pc1 = 0;                   // the length penalty of row i
	int trow = 0;                  // table path: the table row of row i's residue (read a row ahead)
	// profile bytes (profile path, loaded one row ahead) or table values (gathered at the row's start)
	int pf_lo[C], pf_hi[C];
	auto fetch_row = [&](int res) -> int { return (int)tab[res * TAB_W + trow_col]; };
	auto prefetch_tab = [&](int tr) {
#pragma unroll
		for (int k = 0; k < C; ++k)
			pf_lo[k] = __builtin_amdgcn_ds_bpermute(bl[k], tr), pf_hi[k] = __builtin_amdgcn_ds_bpermute(bh[k], tr);
	};
	auto prefetch_prof = [&](int res) {
		const int8_t *pr = prof + res * ROWB + L;
#pragma unroll
		for (int k = 0; k < C; ++k) pf_lo[k] = pr[k * 64], pf_hi[k] = pr[HALF + k * 64];
	};
	if (nl >= 3) {
		const int r0 = sld(rec, 0), r1 = sld(rec, 1);
		d0 = (int)(int8_t)(r0 >> 8), dp1 = (int)(int8_t)(r1 >> 8);
		am1 = (int)(int8_t)(r0 >> 16), a0 = (int)(int8_t)(r1 >> 16);
		rc1 = sld(rec, 2);
		if (nl > 3) rc2 = sld(rec, 3);
		pc1 = pen_of(2);
		if (!PROF) trow = fetch_row(rc1 & 0xff);
		else if (C <= 8) prefetch_prof(rc1 & 0xff);
	}

	int best_log = INT_MIN, best_i = -1, last_i = 1, i = 2, last_ev = NEG;

	auto row = [&](auto Rc) -> bool {
		constexpr int R = decltype(Rc)::value, R2 = (R + 1) % 3, R1 = (R + 2) % 3;
		s2 *H3 = Hs[R], *H2 = Hs[R2], *H1 = Hs[R1];
		int *D3 = dring + R * C * 64 + L;
		const int rc = rc1, rn = rc2, pen_i = pc1;
		rc1 = rn;
		rc2 = i + 2 < nl? sld(rec, i + 2) : 0;
		pc1 = pen_of(i + 1);
		const int res = rc & 0xff;
		dm1 = d0, d0 = dp1, dp1 = (int)(int8_t)(rc >> 8);
		am2 = am1, am1 = a0, a0 = (int)(int8_t)(rc >> 16);
		const s2 GEI = spl(res == 20? fs : ge);
		const s2 DM1 = spl(dm1), D0 = spl(d0), DP1 = spl(dp1), A0 = spl(a0), AM1 = spl(am1), AM2 = spl(am2);

		// this row's scores: the table path gathers them from the table row read during the previous
		// row (and reads the next one); the profile path loaded them during the previous row
		s2 s[C];
		if (!PROF) {
			prefetch_tab(trow);
			if (i + 1 < nl) trow = fetch_row(rn & 0xff);
		} else if (C > 8) prefetch_prof(res);
#pragma unroll
		for (int k = 0; k < C; ++k)   // table path: the low halves of the two words
			s[k] = up((int)__builtin_amdgcn_perm((uint32_t)pf_hi[k], (uint32_t)pf_lo[k], PROF? sel[k] : 0x05040100u));
		if (PROF && C <= 8 && i + 1 < nl) prefetch_prof(rn & 0xff);

		// phase A: every term but the insertion; independent across columns
		const s2 n1 = nh[R1], n2 = nh[R2], n3 = nh[R];
		s2 u[C], m[C], nonI[C];
#pragma unroll
		for (int k = 0; k < C; ++k) u[k] = subs(H1[k], IO), m[k] = vmax(H1[k], H2[k]);
		const s2 u_n = subs(n1, IO), m_n = vmax(n1, n2);
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const s2 h3p = k? H3[k - 1] : n3, up_ = k? u[k - 1] : u_n, mp = k? m[k - 1] : m_n;
			const s2 hm = adds(h3p, s[k]);                                  // H(i-3,j-1) + s
			const s2 dn = subs(vmax(subs(H3[k], GO), up(D3[k * 64])), GEI);   // D
			A[k] = vmax(subs(u[k], DM1), A[k]);                             // phase-0 intron
			B[k] = vmax(subs(up_, D0), B[k]);                               // phase-1
			Cc[k] = vmax(subs(up_, DP1), Cc[k]);                            // phase-2
			const s2 f = subs(vmax(m[k], mp), FS);                          // 4 frameshift moves
			nonI[k] = vmax(vmax(vmax(hm, dn), vmax(subs(A[k], A0), subs(B[k], AM2))), vmax(subs(Cc[k], AM1), f));
			D3[k * 64] = pk(dn);
		}
		// phase B: insertion chain inside the lane (zero carry); opening from an H that is itself an
		// insertion is dominated by extending it, so the chain only needs the non-insertion terms
		s2 I = up(PACKNEG), hh[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			if (k) I = subs(vmax(subs(nonI[k - 1], GO), I), GE);
			hh[k] = vmax(nonI[k], I);
		}
		const s2 X = subs(vmax(subs(nonI[C - 1], GO), I), GE);          // insertion leaving the lane

		// max-plus scan of the lanes' outgoing insertions (biased unsigned, 0 = -32768)
		int S = pk(X) ^ BIAS;
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR1>(S)), DEC1)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR2>(S)), DEC2)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR4>(S)), DEC4)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR8>(S)), DEC8)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_BCAST15>(S)), DEC15)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_BCAST31>(S)), DEC31)));
		const int Tlo = __builtin_amdgcn_readlane(S, 63) & 0xffff;     // leaving the low half (biased)
		const u2 cinu = maxu(upu(dpp0<WAVE_SHR1>(S)), subu(u2{0, (unsigned short)Tlo}, DECL));
		const s2 cin = up(pku(cinu) ^ BIAS);

		// phase C: fold the carry into the cells, masked row maximum
		// (the carry decays by ge per column: two interleaved chains of saturating subtractions, which
		// equal one subtraction of ge * k, so only ge and 2 ge are kept in scalar registers)
		s2 mk[C], ck[2] = {cin, subs(cin, GE)};
		const s2 GE2 = spl(2 * ge);
#pragma unroll
		for (int k = 0; k < C; ++k) {
			if (k >= 2) ck[k & 1] = subs(ck[k & 1], GE2);
			const s2 hk = vmax(hh[k], ck[k & 1]);
			H3[k] = hk;
			mk[k] = LANE_MASK? hk : vmin(hk, lim[k]);
		}
#pragma unroll
		for (int w = 1; w < C; w *= 2)
#pragma unroll
			for (int k = 0; k + w < C; k += 2 * w) mk[k] = vmax(mk[k], mk[k + w]);
		// row maximum over both halves and all lanes
		// (biased by +32768 so that the zero a source-less lane reads is neutral)
		const s2 rm = LANE_MASK? vmin(mk[0], lim[0]) : mk[0];
		int mx = std::max((int)rm.x, (int)rm.y) + 32768;
		mx = std::max(mx, dpp0<ROW_SHR1>(mx));
		mx = std::max(mx, dpp0<ROW_SHR2>(mx));
		mx = std::max(mx, dpp0<ROW_SHR4>(mx));
		mx = std::max(mx, dpp0<ROW_SHR8>(mx));
		mx = std::max(mx, dpp0<ROW_BCAST15>(mx));   // rows 1-3 take the row before's last lane
		mx = std::max(mx, dpp0<ROW_BCAST31>(mx));   // rows 2-3 take lane 31: lane 63 holds the whole wave
		const int rowmax = __builtin_amdgcn_readlane(mx, 63) - 32768;

		// H(i, al-1)
		int ev = 0;
		switch (ke) {
#define NSH_CASE(K) case K: if (K < C) ev = __builtin_amdgcn_readlane(pk(H3[K < C? K : 0]), Le); break;
		NSH_CASE(0) NSH_CASE(1) NSH_CASE(2) NSH_CASE(3) NSH_CASE(4) NSH_CASE(5) NSH_CASE(6) NSH_CASE(7)
		NSH_CASE(8) NSH_CASE(9) NSH_CASE(10) NSH_CASE(11) NSH_CASE(12) NSH_CASE(13) NSH_CASE(14) NSH_CASE(15)
#undef NSH_CASE
		default: break;
		}
		ev = he? hi16(ev) : lo16(ev);
		last_ev = ev;

		// row score, length penalty, best row and x-drop (nasw-sse.c:438-447)
		const int es = ev + eb;
		const int tmp = rowmax > es? rowmax : es;
		const int tlog = tmp - pen_i;
		if (tlog > best_log) {
			best_log = tlog, best_i = i;
#pragma unroll
			for (int k = 0; k < C; ++k) snap[k * 64 + L] = pk(H3[k]);
		}
		// left neighbour of the next row: lane L-1's last column (rotated in); lane 0 gets column -1
		// (neg_inf for rows > 2) in the low half and the low half's last column in the high half
		const int Vr = __builtin_amdgcn_update_dpp(0, pk(H3[C - 1]), WAVE_ROR1, 0xf, 0xf, false);
		nh[R] = up((int)__builtin_amdgcn_perm((uint32_t)Vr, 0x8000u, nsel));
		last_i = i;
		if (best_log - tlog > xdrop) return false;
		return ++i < nl;
	};

	if (nl > 2) {
		// row 2 alone: column -1 holds 0 / -fs for it only, so afterwards lane 0's left neighbour
		// of rows 0 and 1 (slots 1 and 2 when row 2 used slot 0) reads neg_inf in the low half
		bool more = row(std::integral_constant<int, 0>());
		if (L == 0) nh[1].x = NEG, nh[2].x = NEG;
		while (more) {
			if (!row(std::integral_constant<int, 1>())) break;
			if (!row(std::integral_constant<int, 2>())) break;
			if (!row(std::integral_constant<int, 0>())) break;
		}
	}

	// first real column reaching the best row's maximum (nasw-sse.c:448-458)
	int bv = INT_MIN, bj = INT_MAX;
#pragma unroll
	for (int h = 0; h < 2; ++h)
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int j = h * HALF + L * C + k;
			if (j < al) {
				const s2 hm = up(snap[k * 64 + L]);
				const int v = (h? (int)hm.y : (int)hm.x) + (j == al - 1? eb : 0);
				if (v > bv) bv = v, bj = j;
			}
		}
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) {
		const int ov = __shfl_xor(bv, off), oj = __shfl_xor(bj, off);
		if (ov > bv || (ov == bv && oj < bj)) bv = ov, bj = oj;
	}
	{   // every lane holds the same values: all store them (no lane-dependent branch)
		int32_t *o = a.out + 4 * (int64_t)P.out;
		if (P.mode) o[0] = nl, o[1] = al, o[2] = last_ev;   // nasw-sse.c:559: H(last row, al-1)
		else o[0] = best_i + 1, o[1] = bj + 1, o[2] = bv;
		o[3] = last_i - 1;
	}
}

// ---- two problems per wave (matrix scores) ----
// Problem A lives in the low 16-bit halves and problem B in the high halves, each with C columns per
// lane (column c of either problem: lane c / C, slot c % C). The wave's fixed per-row work (the lane
// scan, the row maximum, the neighbour exchange, the loop control) is shared by two problems, and a
// problem's padding granularity is 64 columns instead of 128. Everything that belongs to one problem
// (rows, records, splice signals, length penalty, best row, x-drop) is tracked per half; a problem
// that has stopped leaves its half running on neutral records until the other one stops. The host
// pairs problems of equal C with at least 3 rows each.
DEV int pk2(int lo, int hi) { return (lo & 0xffff) | (hi << 16); }
DEV int sx8(int v) { return (int)(int8_t)(v & 0xff); }

template <int C>
DEV void ext_solve2(const ExtArgs &a, const ExtProb &PA, const ExtProb &PB, int8_t *prof, int *snap, int *dring, const int L)
{
	constexpr bool LANE_MASK = (8 % C) == 0;
	const int go = a.go, ge = a.ge, fs = a.fs, eb = a.eb, xdrop = a.xdrop;
	const int nlA = PA.nl, nlB = PB.nl, alA = PA.al, alB = PB.al;
	const int nlmax = nlA > nlB? nlA : nlB;
	const uint32_t *recA = a.rec + PA.rec_off, *recB = a.rec + PB.rec_off;
	const int plenA = 3 * alA, plenB = 3 * alB;

	// per-lane column constants: row-maximum masks, ds_bpermute sources (B's scores sit in lanes 32..54)
	s2 lim[LANE_MASK? 1 : C];
	int bl[C], bh[C];
	{
		const uint8_t *aasA = a.aas + PA.aa_off, *aasB = a.aas + PB.aa_off;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int c = L * C + k;
			if (!LANE_MASK) lim[k] = s2{(short)(c < PA.npad? 32767 : NEG), (short)(c < PB.npad? 32767 : NEG)};
			bl[k] = (c < alA? aasA[c] : TAB_PAD) * 4;
			bh[k] = (32 + (c < alB? aasB[c] : TAB_PAD)) * 4;
		}
	}
	if (LANE_MASK) lim[0] = s2{(short)(L * C < PA.npad? 32767 : NEG), (short)(L * C < PB.npad? 32767 : NEG)};
	const int trow_col = (L & 31) < TAB_W? (L & 31) : TAB_PAD;
	const bool hiL = L >= 32;
	const lds_u16_t tab = (lds_u16_t)prof;

	const int geC = ge * C;
	const u2 DEC1 = splu(geC), DEC2 = splu(2 * geC), DEC4 = splu(4 * geC), DEC8 = splu(8 * geC);
	const u2 DEC15 = splu(geC * ((L & 15) + 1));
	const u2 DEC31 = splu(L >= 32? geC * (L - 31) : 0xffff);
	const s2 GO = spl(go), GE = spl(ge), FS = spl(fs), GE2 = spl(2 * ge);
	const s2 IO = s2{(short)PA.io, (short)PB.io};

	s2 Hs[3][C], A[C], B[C], Cc[C], nh[3];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		Hs[0][k] = Hs[1][k] = Hs[2][k] = up(PACKNEG);
		dring[(0 * C + k) * 64 + L] = dring[(1 * C + k) * 64 + L] = dring[(2 * C + k) * 64 + L] = PACKNEG;
		A[k] = B[k] = Cc[k] = up(PACKNEG);
		snap[k * 64 + L] = PACKNEG;
	}
	// column -1 of both problems: H(-1,-1) = 0 and H(0,-1) = H(1,-1) = -fs for row 2 only
	nh[0] = nh[1] = nh[2] = up(PACKNEG);
	if (L == 0) nh[0] = s2{0, 0}, nh[1] = nh[2] = s2{(short)-fs, (short)-fs};

	const int LeA = (alA - 1) / C, keA = (alA - 1) % C, LeB = (alB - 1) / C, keB = (alB - 1) % C;

	auto pen_of = [&](int r, int plen) -> int {
		const int d = r - plen;
		return d < 2? 0 : sld(a.pen, d < a.pen_n? d : a.pen_n - 1);
	};
	auto rec_of = [&](const uint32_t *rec, int nl, int r) -> int { return r < nl? sld(rec, r) : 0; };
	auto fetch_rows = [&](int resA, int resB) -> int { return (int)tab[(hiL? resB : resA) * TAB_W + trow_col]; };
	int pf_lo[C], pf_hi[C];
	auto prefetch_tab = [&](int tr) {
#pragma unroll
		for (int k = 0; k < C; ++k)
			pf_lo[k] = __builtin_amdgcn_ds_bpermute(bl[k], tr), pf_hi[k] = __builtin_amdgcn_ds_bpermute(bh[k], tr);
	};

	// splice signals of rows i-1..i+1 as (A, B) pairs; records of rows i and i+1 and the penalty of row i
	int d0p, dp1p, am1p, a0p, dm1p = 0, am2p = 0;
	int rc1A, rc2A, rc1B, rc2B, pc1A, pc1B, trow;
	{
		const int r0A = sld(recA, 0), r1A = sld(recA, 1), r0B = sld(recB, 0), r1B = sld(recB, 1);
		d0p = pk2(sx8(r0A >> 8), sx8(r0B >> 8)), dp1p = pk2(sx8(r1A >> 8), sx8(r1B >> 8));
		am1p = pk2(sx8(r0A >> 16), sx8(r0B >> 16)), a0p = pk2(sx8(r1A >> 16), sx8(r1B >> 16));
		rc1A = sld(recA, 2), rc2A = rec_of(recA, nlA, 3);
		rc1B = sld(recB, 2), rc2B = rec_of(recB, nlB, 3);
		pc1A = pen_of(2, plenA), pc1B = pen_of(2, plenB);
		trow = fetch_rows(rc1A & 0xff, rc1B & 0xff);
	}

	int bestA = INT_MIN, bestiA = -1, lastA = 1, levA = NEG;
	int bestB = INT_MIN, bestiB = -1, lastB = 1, levB = NEG;
	bool liveA = true, liveB = true;
	int i = 2;

	auto row = [&](auto Rc) -> bool {
		constexpr int R = decltype(Rc)::value, R2 = (R + 1) % 3, R1 = (R + 2) % 3;
		s2 *H3 = Hs[R], *H2 = Hs[R2], *H1 = Hs[R1];
		int *D3 = dring + R * C * 64 + L;
		const int rcA = rc1A, rcB = rc1B, rnA = rc2A, rnB = rc2B, penA = pc1A, penB = pc1B;
		rc1A = rc2A, rc2A = rec_of(recA, nlA, i + 2);
		rc1B = rc2B, rc2B = rec_of(recB, nlB, i + 2);
		pc1A = pen_of(i + 1, plenA), pc1B = pen_of(i + 1, plenB);
		const int resA = rcA & 0xff, resB = rcB & 0xff;
		dm1p = d0p, d0p = dp1p, dp1p = pk2(sx8(rcA >> 8), sx8(rcB >> 8));
		am2p = am1p, am1p = a0p, a0p = pk2(sx8(rcA >> 16), sx8(rcB >> 16));
		const s2 GEI = up(pk2(resA == 20? fs : ge, resB == 20? fs : ge));
		const s2 DM1 = up(dm1p), D0 = up(d0p), DP1 = up(dp1p), A0 = up(a0p), AM1 = up(am1p), AM2 = up(am2p);

		// this row's scores from the table rows read during the previous row; then the next rows'
		s2 s[C];
		prefetch_tab(trow);
		if (i + 1 < nlmax) trow = fetch_rows(rnA & 0xff, rnB & 0xff);
#pragma unroll
		for (int k = 0; k < C; ++k) s[k] = up((int)__builtin_amdgcn_perm((uint32_t)pf_hi[k], (uint32_t)pf_lo[k], 0x05040100u));

		// phases A and B: as for one problem (every operation acts on the two halves separately)
		const s2 n1 = nh[R1], n2 = nh[R2], n3 = nh[R];
		s2 u[C], m[C], nonI[C];
#pragma unroll
		for (int k = 0; k < C; ++k) u[k] = subs(H1[k], IO), m[k] = vmax(H1[k], H2[k]);
		const s2 u_n = subs(n1, IO), m_n = vmax(n1, n2);
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const s2 h3p = k? H3[k - 1] : n3, up_ = k? u[k - 1] : u_n, mp = k? m[k - 1] : m_n;
			const s2 hm = adds(h3p, s[k]);
			const s2 dn = subs(vmax(subs(H3[k], GO), up(D3[k * 64])), GEI);
			A[k] = vmax(subs(u[k], DM1), A[k]);
			B[k] = vmax(subs(up_, D0), B[k]);
			Cc[k] = vmax(subs(up_, DP1), Cc[k]);
			const s2 f = subs(vmax(m[k], mp), FS);
			nonI[k] = vmax(vmax(vmax(hm, dn), vmax(subs(A[k], A0), subs(B[k], AM2))), vmax(subs(Cc[k], AM1), f));
			D3[k * 64] = pk(dn);
		}
		s2 I = up(PACKNEG), hh[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			if (k) I = subs(vmax(subs(nonI[k - 1], GO), I), GE);
			hh[k] = vmax(nonI[k], I);
		}
		const s2 X = subs(vmax(subs(nonI[C - 1], GO), I), GE);

		// lane scan of the outgoing insertions, both halves at once; lane 0 takes no carry
		int S = pk(X) ^ BIAS;
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR1>(S)), DEC1)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR2>(S)), DEC2)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR4>(S)), DEC4)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_SHR8>(S)), DEC8)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_BCAST15>(S)), DEC15)));
		S = pku(maxu(upu(S), subu(upu(dpp0<ROW_BCAST31>(S)), DEC31)));
		const s2 cin = up(dpp0<WAVE_SHR1>(S) ^ BIAS);

		s2 mk[C], ck[2] = {cin, subs(cin, GE)};
#pragma unroll
		for (int k = 0; k < C; ++k) {
			if (k >= 2) ck[k & 1] = subs(ck[k & 1], GE2);
			const s2 hk = vmax(hh[k], ck[k & 1]);
			H3[k] = hk;
			mk[k] = LANE_MASK? hk : vmin(hk, lim[k]);
		}
#pragma unroll
		for (int w = 1; w < C; w *= 2)
#pragma unroll
			for (int k = 0; k + w < C; k += 2 * w) mk[k] = vmax(mk[k], mk[k + w]);
		// both problems' row maxima at once, as biased unsigned pairs (a source-less lane reads 0)
		int mx = pk(LANE_MASK? vmin(mk[0], lim[0]) : mk[0]) ^ BIAS;
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_SHR1>(mx))));
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_SHR2>(mx))));
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_SHR4>(mx))));
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_SHR8>(mx))));
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_BCAST15>(mx))));
		mx = pku(maxu(upu(mx), upu(dpp0<ROW_BCAST31>(mx))));
		const int rmx = __builtin_amdgcn_readlane(mx, 63);

		// H(i, al-1) of each problem
		int evA = 0, evB = 0;
		switch (keA) {
#define NSH_CASE(K) case K: if (K < C) evA = __builtin_amdgcn_readlane(pk(H3[K < C? K : 0]), LeA); break;
		NSH_CASE(0) NSH_CASE(1) NSH_CASE(2) NSH_CASE(3)
#undef NSH_CASE
		default: break;
		}
		switch (keB) {
#define NSH_CASE(K) case K: if (K < C) evB = __builtin_amdgcn_readlane(pk(H3[K < C? K : 0]), LeB); break;
		NSH_CASE(0) NSH_CASE(1) NSH_CASE(2) NSH_CASE(3)
#undef NSH_CASE
		default: break;
		}
		evA = lo16(evA), evB = hi16(evB);

		// row score, length penalty, best row and x-drop of each live problem (nasw-sse.c:438-447)
		if (liveA) {
			const int rowmax = (rmx & 0xffff) - 32768, es = evA + eb;
			const int tlog = (rowmax > es? rowmax : es) - penA;
			levA = evA;
			if (tlog > bestA) {
				bestA = tlog, bestiA = i;
#pragma unroll
				for (int k = 0; k < C; ++k) ((short*)snap)[2 * (k * 64 + L)] = H3[k].x;
			}
			lastA = i;
			if (bestA - tlog > xdrop || i + 1 >= nlA) liveA = false;
		}
		if (liveB) {
			const int rowmax = (int)((unsigned)rmx >> 16) - 32768, es = evB + eb;
			const int tlog = (rowmax > es? rowmax : es) - penB;
			levB = evB;
			if (tlog > bestB) {
				bestB = tlog, bestiB = i;
#pragma unroll
				for (int k = 0; k < C; ++k) ((short*)snap)[2 * (k * 64 + L) + 1] = H3[k].y;
			}
			lastB = i;
			if (bestB - tlog > xdrop || i + 1 >= nlB) liveB = false;
		}
		// left neighbour of the next row: lane L-1's last column; lane 0 keeps column -1 (neg_inf)
		nh[R] = up(__builtin_amdgcn_update_dpp(PACKNEG, pk(H3[C - 1]), WAVE_SHR1, 0xf, 0xf, false));
		++i;
		return liveA || liveB;
	};

	bool more = row(std::integral_constant<int, 0>());
	if (L == 0) nh[1] = nh[2] = up(PACKNEG);
	while (more) {
		if (!row(std::integral_constant<int, 1>())) break;
		if (!row(std::integral_constant<int, 2>())) break;
		if (!row(std::integral_constant<int, 0>())) break;
	}

	// each problem's first real column reaching its best row's maximum (nasw-sse.c:448-458)
#pragma unroll
	for (int h = 0; h < 2; ++h) {
		const ExtProb &P = h? PB : PA;
		const int al = P.al;
		int bv = INT_MIN, bj = INT_MAX;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int j = L * C + k;
			if (j < al) {
				const s2 hm = up(snap[k * 64 + L]);
				const int v = (h? (int)hm.y : (int)hm.x) + (j == al - 1? eb : 0);
				if (v > bv) bv = v, bj = j;
			}
		}
#pragma unroll
		for (int off = 32; off >= 1; off >>= 1) {
			const int ov = __shfl_xor(bv, off), oj = __shfl_xor(bj, off);
			if (ov > bv || (ov == bv && oj < bj)) bv = ov, bj = oj;
		}
		int32_t *o = a.out + 4 * (int64_t)P.out;
		if (P.mode) o[0] = P.nl, o[1] = al, o[2] = h? levB : levA;
		else o[0] = (h? bestiB : bestiA) + 1, o[1] = bj + 1, o[2] = bv;
		o[3] = (h? lastB : lastA) - 1;
	}
}

// persistent waves over one group's queue; CLO..CHI columns per lane per half. The waves-per-SIMD
// target trades registers for latency hiding (measured: 6 / 3 / 1 for C <= 4 / <= 8 / <= 16; the
// merged queue of C 1..8 takes the C <= 8 budget).
template <int CLO, int CHI, bool PROF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CHI <= 4? NSH_EXT_WPE0 : CLO == 1 && CHI == 8? NSH_EXT_WPE01 : CHI <= 8? NSH_EXT_WPE1 : 1)))
ext_dp_kernel(ExtArgs a)
{
	__shared__ __attribute__((aligned(16))) int8_t prof[PROF? NRES * 128 * CHI : NRES * TAB_W * 2];
	__shared__ int snap[64 * CHI], dring[3 * 64 * CHI];
	const int L = threadIdx.x;
	if (!PROF) {   // the matrix as an int16 [row residue][column residue] table, padding column -32768
		int16_t *tab = (int16_t*)prof;
		for (int t = L; t < NRES * TAB_W; t += 64) {
			const int r = t / TAB_W, aa = t % TAB_W;
			tab[t] = (int16_t)(aa < TAB_PAD? a.mat[r * a.asize + aa] : NEG);
		}
		__syncthreads();
	}
	for (;;) {
		// ticket: every lane takes part in the atomic (lane 0 adds 1, the others 0), so the loop has
		// no lane-dependent branch; the ticket is lane 0's value, wave-uniform
		const int p = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, L == 0? 1 : 0));
		if (p >= a.n) break;
		const ExtProb P = a.prob[p];
		if (!PROF && CLO <= 4 && P.pair >= 0) {   // two problems of P.C columns per lane each
			const ExtProb Q = a.pbase[P.pair];
			if ((int64_t)(P.nl > Q.nl? P.nl : Q.nl) * P.C >= kLongWork) __builtin_amdgcn_s_setprio(2);
			else __builtin_amdgcn_s_setprio(0);
#define NSH_C2(K) if (!PROF && K >= CLO && K <= CHI && P.C == K) { ext_solve2<(K >= CLO && K <= CHI)? K : CLO>(a, P, Q, prof, snap, dring, L); continue; }
			NSH_C2(1) NSH_C2(2) NSH_C2(3) NSH_C2(4)
#undef NSH_C2
			continue;
		}
		const int C = P.C < CLO? CLO : P.C > CHI? CHI : P.C;
		// a problem of many rows (an extension over a window of tens of kilobases) sets the length of
		// the whole launch: its wave gets the issue priority over the waves of short problems sharing
		// its SIMD, which fill the cycles it leaves
		if ((int64_t)P.nl * C >= kLongWork) __builtin_amdgcn_s_setprio(2);
		else __builtin_amdgcn_s_setprio(0);
#define NSH_C(K) if (K >= CLO && K <= CHI && C == K) { ext_solve<(K >= CLO && K <= CHI)? K : CLO, PROF>(a, P, prof, snap, dring, L); continue; }
		NSH_C(1) NSH_C(2) NSH_C(3) NSH_C(4) NSH_C(5) NSH_C(6) NSH_C(7) NSH_C(8)
		NSH_C(9) NSH_C(10) NSH_C(11) NSH_C(12) NSH_C(13) NSH_C(14) NSH_C(15) NSH_C(16)
#undef NSH_C
	}
}

} // namespace nsh_ext
#endif
