// Device side of the work-group-per-problem spliced DP (global_dp.hip describes the engine).
// Included by the global_dp_k*.hip translation units, one per arithmetic / mode.
//
// The reference's trace words depend on its striped SIMD evaluation order: with vsize lanes
// (8 for int16, 4 for int32) and slen = ceil(al/vsize), column c belongs to segment c/slen,
// the first pass restricts the insertion to the segment (reset at its start), and a lazy-F pass
// carries insertions across segments (nasw-sse.c:461-567, 861-973). This kernel evaluates the
// same quantities row by row:
//   * phase A, per cell, independent across columns: the nine non-insertion candidates, their
//     first strict maximum (order D, A, B, C, F1, F2, G1, G2 after M) and the extension bits 5-8;
//   * the segment-local first-pass insertion I_fp (reset at every segment start) and the exact
//     insertion I_ex, each a chain inside the lane plus a max-plus scan across lanes and waves
//     (the I_fp scan is segmented: a lane holding a segment start stops the carry);
//   * H_fp = max(non-insertion, I_fp) gives bits 0-4 exactly as the first pass does; the lazy-F
//     pass raises a cell iff I_ex > H_fp (bit 9 for int16; for int32 the word is overwritten
//     with 1 | ext, ext taken from the segment that delivered the carry), and H = max(., I_ex).
// A work-group of NW <= 16 waves solves one problem (NW*64 lanes, C <= 8 columns per lane); waves
// exchange their scan totals and boundary columns through LDS with two barriers per row.
// Values live in 32-bit registers with the reference's arithmetic (int16 saturating or int32
// wrapping). Trace words go to HBM, one uint16 per real cell.
//
// EXT mode (int16, no trace) is the score-only extension DP of nasw-sse.c:360-459 for proteins
// too long for the one-wave extension kernel: columns up to the reference's padded width npad
// take part (padded columns score -32768 on the diagonal move), each row's maximum over them and
// H(i, al-1) + end_bonus give the row score, the length penalty comes from a host table, the
// best row's H values are kept, and the DP stops once the best row score exceeds the current one
// by more than x-drop. The row's decision is taken right after the next row's first barrier,
// from per-wave maxima in LDS, so it costs no extra barrier.
#ifndef NSH_GLOBAL_DP_IMPL_H
#define NSH_GLOBAL_DP_IMPL_H
#include <hip/hip_runtime.h>
#include <climits>
#include <type_traits>
#include "global_dp_types.h"

namespace nsh_glob {

#define DEV static __device__ __forceinline__

// arithmetic of the kernel being matched
struct Arith16 {
	static constexpr int NEG = -32768;
	static constexpr bool IS32 = false;
	DEV int add(int a, int b) { return min(max(a + b, -32768), 32767); }
	DEV int sub(int a, int b) { return min(max(a - b, -32768), 32767); }
};
struct Arith32 {
	static constexpr int NEG = -0x40000000;
	static constexpr bool IS32 = true;
	DEV int add(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
	DEV int sub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }
};

enum : int { ROW_SHR1 = 0x111, ROW_SHR2 = 0x112, ROW_SHR4 = 0x114, ROW_SHR8 = 0x118,
	WAVE_SHR1 = 0x138, ROW_BCAST15 = 0x142, ROW_BCAST31 = 0x143 };
template <int CTRL> DEV int dppk(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false); }

// LDS exchange area of a work-group (besides the profile)
struct Xch {
	int nbr[2][kNWMax];            // H of the wave's last column, by row parity
	int lfp[2][kNWMax][2];         // I_fp, H_fp of the wave's last column, by row parity
	int tex[kNWMax], tfp[kNWMax], ffp[kNWMax];
	int seg[8];                    // int32: O and flag of segments 0..2
	int rmax[2][kNWMax];           // EXT: each wave's row maximum, by row parity
	int ev[2];                     // EXT: H(i, al-1), by row parity
	int bv[kNWMax], bj[kNWMax];    // EXT: each wave's best column of the best row
};

// in-wave max-plus scan, decay d per lane: S(L) = max_{L'<=L} X(L') - d (L - L')
template <class AR>
DEV int scan_maxplus(int S, int d, int L)
{
	S = max(S, AR::sub(dppk<ROW_SHR1>(S), d));
	S = max(S, AR::sub(dppk<ROW_SHR2>(S), 2 * d));
	S = max(S, AR::sub(dppk<ROW_SHR4>(S), 4 * d));
	S = max(S, AR::sub(dppk<ROW_SHR8>(S), 8 * d));
	{ const int t = dppk<ROW_BCAST15>(S); if (L & 16) S = max(S, AR::sub(t, d * ((L & 15) + 1))); }
	{ const int t = dppk<ROW_BCAST31>(S); if (L >= 32) S = max(S, AR::sub(t, d * (L - 31))); }
	return S;
}
// segmented: F(L) = 1 when lane L holds a segment start; the carry does not cross it. On return
// F(L) = 1 when some lane in [0, L] holds one.
template <class AR>
DEV int scan_maxplus_seg(int S, int &F, int d, int L)
{
#define NSH_SEG_STEP(CTRL, DIST, COND) { \
		const int t = dppk<CTRL>(S), tf = dppk<CTRL>(F); \
		if ((COND) && !F) S = max(S, AR::sub(t, (DIST))); \
		if (COND) F |= tf; }
	NSH_SEG_STEP(ROW_SHR1, d, (L & 15) >= 1)
	NSH_SEG_STEP(ROW_SHR2, 2 * d, (L & 15) >= 2)
	NSH_SEG_STEP(ROW_SHR4, 4 * d, (L & 15) >= 4)
	NSH_SEG_STEP(ROW_SHR8, 8 * d, (L & 15) >= 8)
	NSH_SEG_STEP(ROW_BCAST15, d * ((L & 15) + 1), (L & 16) != 0)
	NSH_SEG_STEP(ROW_BCAST31, d * (L - 31), L >= 32)
#undef NSH_SEG_STEP
	return S;
}

DEV int wave_max(int v)
{
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
	return v;
}

template <class AR, int C, bool TRACE, bool EXT>
DEV void glob_solve(const GlobArgs &a, const GlobProb &P, int8_t *prof, Xch &X)
{
	static_assert(!EXT || (!TRACE && !AR::IS32), "the extension mode is score-only int16");
	const int NEG = AR::NEG;
	const int L = threadIdx.x & 63, w = threadIdx.x >> 6, NW = P.NW, G = threadIdx.x;
	const int nl = P.nl, al = P.al, slen = P.slen, asz = a.asize;
	const int go = a.go, ge = a.ge, fs = a.fs, io = P.io, goe = go + ge;
	const int ncol = EXT? P.npad : al;           // columns that take part in the DP
	const uint8_t *aas = a.aas + P.aa_off;
	const int8_t *pssm = P.pssm_off >= 0? a.pssm + P.pssm_off : nullptr;
	const uint32_t *rec = a.rec + P.rec_off;
	const int W = 64 * C * NW;                   // columns covered by the work-group
	// column constants
	int jpos[C];
	bool valid[C], real[C];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		const int c = G * C + k;
		valid[k] = c < ncol, real[k] = c < al;
		jpos[k] = c % slen;
	}
	int lane_reset = 0;
	if (!EXT) {
#pragma unroll
		for (int k = 0; k < C; ++k) lane_reset |= valid[k] && jpos[k] == 0;
	}
	// scores: up to 4 columns per lane an int8 profile in LDS, prof[r * W + c]; with more columns
	// (proteins over 4096 columns) the profile would not fit the LDS, so the 22 x 22 matrix is
	// staged there instead and looked up with the lane's residue codes, or PSSM rows are read
	// from memory
	constexpr bool BIG = C > 4;
	int aak[BIG? C : 1];
	int64_t qrow[BIG? C : 1];
	if (!BIG) {
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int c = G * C + k;
			const int aa = real[k]? aas[c] : 0;
			for (int r = 0; r < 22; ++r)
				prof[r * W + c] = real[k]? (pssm? pssm[(P.left? al - 1 - c : c) * asz + r] : a.mat[r * asz + aa]) : 0;
		}
	} else {
		for (int t = G; t < 22 * 22; t += 64 * NW) prof[t] = a.mat[(t / 22) * asz + t % 22];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int c = G * C + k;
			aak[k] = real[k]? aas[c] : 0;
			qrow[k] = real[k]? (int64_t)(P.left? al - 1 - c : c) * asz : 0;
		}
	}
	__syncthreads();

	int Hs[3][C], Ds[3][C], A[C], B[C], Cc[C], nh[3];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		Hs[0][k] = Hs[1][k] = Hs[2][k] = Ds[0][k] = Ds[1][k] = Ds[2][k] = NEG;
		A[k] = B[k] = Cc[k] = NEG;
	}
	// left neighbour of the lane's first column, rows i-3 (nh[0]), i-2 (nh[1]), i-1 (nh[2]) at row 2:
	// column -1 holds H(-1,-1)=0 and H(0,-1)=H(1,-1)=-fs when row 2 is computed
	nh[0] = nh[1] = nh[2] = NEG;
	if (G == 0) nh[0] = 0, nh[1] = nh[2] = -fs;

	int recv = 0, recn = 0;
	auto load_blk = [&](int i0) { const int i = i0 + L; return i < nl? (int)rec[i] : 0; };
	recv = load_blk(0);
	if (nl > 64) recn = load_blk(64);
	int d0 = 0, dp1 = 0, am1 = 0, a0 = 0, dm1 = 0, am2 = 0;
	if (nl >= 3) {
		const int r0 = __builtin_amdgcn_readlane(recv, 0), r1 = __builtin_amdgcn_readlane(recv, 1);
		d0 = (int)(int8_t)(r0 >> 8), dp1 = (int)(int8_t)(r1 >> 8);
		am1 = (int)(int8_t)(r0 >> 16), a0 = (int)(int8_t)(r1 >> 16);
	}
	// the I_fp a segment starts with: I = last_h = neg_inf -> sub(max(sub(NEG, go), NEG), ge)
	const int fp_start = AR::sub(max(AR::sub(NEG, go), NEG), ge);
	const int geC = ge * C, geW = ge * C * 64;
	uint16_t *tr = TRACE? a.trace + P.tr_off : nullptr;
	// the first cell of lane 0 of waves w > 0 is finished one row later (its bit 4 needs the
	// previous wave's last column): pending word, its fire flag and column
	int z0_pend = 0, f0_pend = 0, r0_pend = 0;
	int i = 2;
	// EXT: row scores, best row, its H values (snap) and x-drop
	const int pen_len = 3 * al, eb = a.eb;
	const int ge_ = (al - 1) / C, ke = (al - 1) % C;   // owner of column al-1
	int snap[EXT? C : 1];
#pragma unroll
	for (int k = 0; k < (EXT? C : 1); ++k) snap[k] = NEG;
	int best_log = INT_MIN, best_i = -1, last_i = 1, last_ev = NEG;
	// the decision of row q, from the per-wave maxima written during row q; Hq: its H values.
	// Returns false when the x-drop stops the DP after row q.
	auto decide = [&](int q, const int *Hq) -> bool {
		int rowmax = X.rmax[q & 1][0];
		for (int v = 1; v < NW; ++v) rowmax = max(rowmax, X.rmax[q & 1][v]);
		const int ev = X.ev[q & 1];
		const int d = q - pen_len;
		const int pen = d < 2? 0 : a.pen[d < a.pen_n? d : a.pen_n - 1];
		const int es = ev + eb;
		const int tlog = (rowmax > es? rowmax : es) - pen;
		if (tlog > best_log) {
			best_log = tlog, best_i = q;
#pragma unroll
			for (int k = 0; k < C; ++k) snap[k] = Hq[k];
		}
		last_i = q, last_ev = ev;
		return best_log - tlog <= a.xdrop;
	};

	auto row = [&](auto Rc) -> bool {
		constexpr int R = decltype(Rc)::value, R2 = (R + 1) % 3, R1 = (R + 2) % 3;
		int *H3 = Hs[R], *H2 = Hs[R2], *H1 = Hs[R1], *D3 = Ds[R];
		__syncthreads();   // B1: last row's boundary values (and, EXT, its row maxima) are in LDS
		if (EXT && i > 2 && !decide(i - 1, H1)) return false;
		if (L == 0 && w > 0 && i > 2) {
			nh[R1] = X.nbr[(i - 1) & 1][w - 1];
			// finish the previous row's first word of this wave
			if (TRACE) {
				const int ipl = X.lfp[(i - 1) & 1][w - 1][0], hpl = X.lfp[(i - 1) & 1][w - 1][1];
				const bool reset0 = valid[0] && jpos[0] == 0;
				if (!(AR::IS32 && f0_pend) && !reset0 && ipl > AR::sub(hpl, go)) z0_pend |= 1 << 4;
				if (valid[0]) tr[(int64_t)r0_pend * al + G * C] = (uint16_t)z0_pend;
			}
		}
		if ((i & 63) == 0) { recv = recn; if (i + 64 < nl) recn = load_blk(i + 64); }
		const int rc = __builtin_amdgcn_readlane(recv, i & 63);
		const int res = rc & 0xff;
		dm1 = d0, d0 = dp1, dp1 = (int)(int8_t)(rc >> 8);
		am2 = am1, am1 = a0, a0 = (int)(int8_t)(rc >> 16);
		const int gei = res == 20? fs : ge;
		const int8_t *pr = prof + res * W + G * C;
		int sc[C];
#pragma unroll
		for (int k = 0; k < C; ++k)
			sc[k] = !BIG? (int)pr[k] : pssm? (int)pssm[qrow[k] + res] : (int)prof[res * 22 + aak[k]];
		const int n1 = nh[R1], n2 = nh[R2], n3 = nh[R];

		// phase A
		int Mv[C], rest[C], yrest[C], bits[C], tgo[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int h3p = k? H3[k - 1] : n3, h1p = k? H1[k - 1] : n1, h2p = k? H2[k - 1] : n2;
			int zz = 0, t, u;
			// padded columns of the extension mode score neg_inf on the diagonal move
			Mv[k] = AR::add(h3p, (EXT && !real[k])? NEG : sc[k]);
			u = AR::sub(H3[k], go);
			if (TRACE && D3[k] > u) zz |= 1 << 5;
			const int dn = AR::sub(max(u, D3[k]), gei);
			t = AR::sub(AR::sub(H1[k], io), dm1);
			if (TRACE && A[k] > t) zz |= 1 << 6;
			A[k] = max(t, A[k]);
			const int av = AR::sub(A[k], a0);
			u = AR::sub(h1p, io);
			t = AR::sub(u, d0);
			if (TRACE && B[k] > t) zz |= 1 << 7;
			B[k] = max(t, B[k]);
			const int bv = AR::sub(B[k], am2);
			t = AR::sub(u, dp1);
			if (TRACE && Cc[k] > t) zz |= 1 << 8;
			Cc[k] = max(t, Cc[k]);
			const int cv = AR::sub(Cc[k], am1);
			const int f6 = AR::sub(H1[k], fs), f7 = AR::sub(H2[k], fs), f8 = AR::sub(h1p, fs), f9 = AR::sub(h2p, fs);
			int best = dn, y = 2;
			if (av > best) best = av, y = 3;
			if (bv > best) best = bv, y = 4;
			if (cv > best) best = cv, y = 5;
			if (f6 > best) best = f6, y = 6;
			if (f7 > best) best = f7, y = 7;
			if (f8 > best) best = f8, y = 8;
			if (f9 > best) best = f9, y = 9;
			rest[k] = best, yrest[k] = y, bits[k] = zz;
			tgo[k] = AR::sub(max(Mv[k], best), go);
			D3[k] = dn;
		}
		// local chains with no carry, the lane's outgoing insertions
		int Iex = NEG, Ifp = NEG;
#pragma unroll
		for (int k = 1; k < C; ++k) {
			Iex = AR::sub(max(tgo[k - 1], Iex), ge);
			if (!EXT) Ifp = (valid[k] && jpos[k] == 0)? fp_start : AR::sub(max(tgo[k - 1], Ifp), ge);
		}
		int Sex = scan_maxplus<AR>(AR::sub(max(tgo[C - 1], Iex), ge), geC, L);
		int Fl = lane_reset, Sfp = NEG;
		if (!EXT) Sfp = scan_maxplus_seg<AR>(AR::sub(max(tgo[C - 1], Ifp), ge), Fl, geC, L);
		if (L == 63) X.tex[w] = Sex, X.tfp[w] = Sfp, X.ffp[w] = Fl;
		__syncthreads();   // B2: wave totals
		// carries entering this wave (the column before it; column -1 for wave 0)
		int cw_ex = fp_start, cw_fp = fp_start;
		for (int v = 0; v < w; ++v) {
			cw_ex = max(X.tex[v], AR::sub(cw_ex, geW));
			if (!EXT) cw_fp = X.ffp[v]? X.tfp[v] : max(X.tfp[v], AR::sub(cw_fp, geW));
		}
		int cin_ex = dppk<WAVE_SHR1>(Sex), cin_fp = 0, fx = 0;
		if (!EXT) cin_fp = dppk<WAVE_SHR1>(Sfp), fx = dppk<WAVE_SHR1>(Fl);
		if (L == 0) cin_ex = cw_ex, cin_fp = cw_fp;
		else {
			cin_ex = max(cin_ex, AR::sub(cw_ex, geC * L));
			if (!EXT && !fx) cin_fp = max(cin_fp, AR::sub(cw_fp, geC * L));
		}
		if (AR::IS32) {
			// first-pass outgoing of segments 0..2 for the segment-start words of lazy-F
			int ifp = cin_fp;
#pragma unroll
			for (int k = 0; k < C; ++k) {
				if (k) ifp = AR::sub(max(tgo[k - 1], ifp), ge);
				if (valid[k] && jpos[k] == 0) ifp = fp_start;
				const int c = G * C + k;
				if (valid[k] && jpos[k] == slen - 1 && c / slen < 3) {
					const int hfp = max(Mv[k], max(ifp, rest[k]));
					const int ho = AR::sub(hfp, goe), ie = AR::sub(ifp, ge);
					X.seg[2 * (c / slen)] = max(ho, ie), X.seg[2 * (c / slen) + 1] = ie > ho;
				}
			}
			__syncthreads();
		}
		// final pass
		int z[C], Hf[C], fires0 = 0;
		int iex = cin_ex, ifp = cin_fp, hfp_prev = NEG, ifp_prev = NEG;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			if (k) {
				iex = AR::sub(max(tgo[k - 1], iex), ge);
				if (!EXT) ifp = AR::sub(max(tgo[k - 1], ifp), ge);
			}
			const int m = Mv[k], r = rest[k];
			if (EXT) {
				Hf[k] = max(max(m, r), iex);
				z[k] = 0;
				continue;
			}
			const bool reset = valid[k] && jpos[k] == 0;
			if (reset) ifp = fp_start;
			int zz = bits[k];
			if (reset) { if (AR::IS32) zz |= 1 << 4; }
			else if (k && ifp_prev > AR::sub(hfp_prev, go)) zz |= 1 << 4;
			const int hfp = max(m, max(ifp, r));
			zz |= (m >= ifp && m >= r)? 0 : (ifp >= r? 1 : yrest[k]);
			const bool fire = iex > hfp;
			if (!AR::IS32) { if (fire) zz |= 1 << 9; }
			else if (fire) {
				int ext = 16;
				if (reset) {   // segment start: which lazy-F iteration delivered the carry
					const int sm = (G * C + k) / slen;
					int F = NEG;
					for (int q = 2; q <= sm; ++q) F = AR::sub(max(q == 2? NEG : F, X.seg[2 * (q - 2)]), ge * slen);
					ext = (sm == 1 || X.seg[2 * (sm - 1)] >= F)? (X.seg[2 * (sm - 1) + 1]? 16 : 0) : 16;
				}
				zz = 1 | ext;
			}
			if (k == 0) fires0 = fire;
			z[k] = zz;
			Hf[k] = max(hfp, iex);
			hfp_prev = hfp, ifp_prev = ifp;
		}
		if (TRACE) {
			// bit 4 of slot 0: previous lane's last column (inside the wave)
			const int ipl = dppk<WAVE_SHR1>(ifp_prev), hpl = dppk<WAVE_SHR1>(hfp_prev);
			const bool reset0 = valid[0] && jpos[0] == 0;
			if (L > 0 && !(AR::IS32 && fires0) && !reset0 && ipl > AR::sub(hpl, go)) z[0] |= 1 << 4;
			uint16_t *dst = tr + (int64_t)i * al + G * C;
#pragma unroll
			for (int k = 0; k < C; ++k)
				if (valid[k] && (k || L || w == 0)) dst[k] = (uint16_t)z[k];
			if (L == 0 && w > 0) z0_pend = z[0], f0_pend = fires0, r0_pend = i;
		}
		if (L == 63) {
			X.nbr[i & 1][w] = Hf[C - 1];
			if (TRACE) X.lfp[i & 1][w][0] = ifp_prev, X.lfp[i & 1][w][1] = hfp_prev;
		}
#pragma unroll
		for (int k = 0; k < C; ++k) H3[k] = Hf[k];
		if (EXT) {   // this row's maximum over the reference's columns, and H(i, al-1)
			int mx = NEG;
#pragma unroll
			for (int k = 0; k < C; ++k) if (valid[k]) mx = max(mx, Hf[k]);
			mx = wave_max(mx);
			if (L == 0) X.rmax[i & 1][w] = mx;
#pragma unroll
			for (int k = 0; k < C; ++k) if (k == ke && G == ge_) X.ev[i & 1] = Hf[k];
		}
		const int Vs = dppk<WAVE_SHR1>(Hf[C - 1]);
		nh[R] = L == 0? NEG : Vs;                 // lane 0 of waves > 0 is filled from LDS next row
		if (i == 2 && G == 0) nh[R2] = NEG, nh[R1] = NEG;
		++i;
		return true;
	};
	bool stopped = false;
	if (nl > 2) {
		for (;;) {
			if (!row(std::integral_constant<int, 0>())) { stopped = true; break; }
			if (i >= nl) break;
			if (!row(std::integral_constant<int, 1>())) { stopped = true; break; }
			if (i >= nl) break;
			if (!row(std::integral_constant<int, 2>())) { stopped = true; break; }
			if (i >= nl) break;
		}
		__syncthreads();
		if (EXT && !stopped) {   // the last row's decision
			const int sl = (nl - 3) % 3;
			int Hl[C];
#pragma unroll
			for (int k = 0; k < C; ++k) Hl[k] = sl == 0? Hs[0][k] : sl == 1? Hs[1][k] : Hs[2][k];
			(void)decide(nl - 1, Hl);
		}
		if (TRACE && L == 0 && w > 0) {
			const int ipl = X.lfp[(i - 1) & 1][w - 1][0], hpl = X.lfp[(i - 1) & 1][w - 1][1];
			const bool reset0 = valid[0] && jpos[0] == 0;
			if (!(AR::IS32 && f0_pend) && !reset0 && ipl > AR::sub(hpl, go)) z0_pend |= 1 << 4;
			if (valid[0]) tr[(int64_t)r0_pend * al + G * C] = (uint16_t)z0_pend;
		}
	}
	if (EXT) {
		// first real column reaching the best row's maximum (nasw-sse.c:448-458)
		int bv = INT_MIN, bj = INT_MAX;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int j = G * C + k;
			if (real[k]) {
				const int v = snap[k] + (j == al - 1? eb : 0);
				if (v > bv) bv = v, bj = j;
			}
		}
#pragma unroll
		for (int off = 32; off >= 1; off >>= 1) {
			const int ov = __shfl_xor(bv, off), oj = __shfl_xor(bj, off);
			if (ov > bv || (ov == bv && oj < bj)) bv = ov, bj = oj;
		}
		if (L == 0) X.bv[w] = bv, X.bj[w] = bj;
		__syncthreads();
		if (G == 0) {
			for (int v = 1; v < NW; ++v)
				if (X.bv[v] > bv || (X.bv[v] == bv && X.bj[v] < bj)) bv = X.bv[v], bj = X.bj[v];
			int32_t *o = a.ext_out + 4 * (int64_t)P.out;
			if (P.mode) o[0] = nl, o[1] = al, o[2] = last_ev;   // nasw-sse.c:559: H(last row, al-1)
			else o[0] = best_i + 1, o[1] = bj + 1, o[2] = bv;
			o[3] = last_i - 1;
		}
		return;
	}
	// score = H(nl-1, al-1); the last row computed sits in Hs[(nl-3) % 3]
	const int last = nl > 2? (nl - 3) % 3 : -1;
	if (G == ge_) {
		int sc = NEG;
#pragma unroll
		for (int k = 0; k < C; ++k)
			if (k == ke) sc = last == 0? Hs[0][k] : last == 1? Hs[1][k] : last == 2? Hs[2][k] : NEG;
		a.score[P.out] = sc;
	}
}

// one work-group per problem. MAXT: the work-group size bound (512 keeps the column counts up to
// 4 free of spills; 1024 serves proteins longer than 2048, up to 8 columns per lane). CK > 0: a
// kernel for one-wave problems of exactly CK columns per lane only, whose registers are sized for
// that case alone (the general kernel's 178 VGPRs allow two waves per SIMD, and one-wave problems
// of a few hundred cells are latency-bound)
template <class AR, bool TRACE, bool EXT, int MAXT, int CK = 0>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(CK == 1? 7 : CK == 2? 5 : 1))) glob_kernel(GlobArgs a)
{
	extern __shared__ int8_t prof[];
	__shared__ Xch X;
	const GlobProb P = a.prob[blockIdx.x];
	// a problem of many rows sets the end of its bucket's launch (its rows are dependent): it takes
	// the issue priority over the short problems sharing its SIMD, which fill the cycles it leaves
	if (P.nl >= kLongRowsG) __builtin_amdgcn_s_setprio(2);
	if constexpr (CK > 0) {
		glob_solve<AR, CK, TRACE, EXT>(a, P, prof, X);
		return;
	}
	switch (P.C) {
	case 1: glob_solve<AR, 1, TRACE, EXT>(a, P, prof, X); break;
	case 2: glob_solve<AR, 2, TRACE, EXT>(a, P, prof, X); break;
	case 3: glob_solve<AR, 3, TRACE, EXT>(a, P, prof, X); break;
	case 4: glob_solve<AR, 4, TRACE, EXT>(a, P, prof, X); break;
	default:
		if constexpr (MAXT > 512) {
			switch (P.C) {
			case 5: glob_solve<AR, 5, TRACE, EXT>(a, P, prof, X); break;
			case 6: glob_solve<AR, 6, TRACE, EXT>(a, P, prof, X); break;
			case 7: glob_solve<AR, 7, TRACE, EXT>(a, P, prof, X); break;
			default: glob_solve<AR, 8, TRACE, EXT>(a, P, prof, X); break;
			}
		}
		break;
	}
}

// launches one bucket of problems that share NW (grid = n problems, 64 * NW threads each)
template <class AR, bool TRACE, bool EXT>
void launch_bucket(const GlobArgs &a, int n, int NW, int C, hipStream_t s)
{
	const size_t lds = C <= 4? (size_t)22 * 64 * C * NW : 512;   // the profile, or the matrix (C > 4)
	if (NW == 1 && C == 1) hipLaunchKernelGGL((glob_kernel<AR, TRACE, EXT, 64, 1>), dim3(n), dim3(64), lds, s, a);
	else if (NW == 1 && C == 2) hipLaunchKernelGGL((glob_kernel<AR, TRACE, EXT, 64, 2>), dim3(n), dim3(64), lds, s, a);
	else if (NW <= 8 && C <= 4) hipLaunchKernelGGL((glob_kernel<AR, TRACE, EXT, 512>), dim3(n), dim3(64 * NW), lds, s, a);
	else hipLaunchKernelGGL((glob_kernel<AR, TRACE, EXT, 1024>), dim3(n), dim3(64 * NW), lds, s, a);
}

template <class AR, bool TRACE, bool EXT>
hipError_t set_lds_limit()
{
	const int big = 22 * 64 * 4 * kNWMax;   // profile bytes of the largest work-group that stages one
	return hipFuncSetAttribute(reinterpret_cast<const void*>(&glob_kernel<AR, TRACE, EXT, 1024>),
			hipFuncAttributeMaxDynamicSharedMemorySize, big);
}

} // namespace nsh_glob
#endif
