// Global spliced DP with CIGAR on gfx950 (ns_global_gs16b / ns_global_gs32b:
// nasw-sse.c:351-567 and 779-973; backtrack nasw-sse.c:44-93).
//
// The reference's trace words depend on its striped SIMD evaluation order: with vsize lanes
// (8 for int16, 4 for int32) and slen = ceil(al/vsize), column c belongs to segment c/slen,
// the first pass restricts the insertion to the segment (reset at its start), and a lazy-F pass
// carries insertions across segments. This kernel evaluates the same quantities row by row:
//   * phase A, per cell, independent across columns: the nine non-insertion candidates, their
//     first strict maximum (order D, A, B, C, F1, F2, G1, G2 after M) and the extension bits 5-8;
//   * the segment-local first-pass insertion I_fp (reset at every segment start) and the exact
//     insertion I_ex, each a chain inside the lane plus a max-plus scan across lanes and waves
//     (the I_fp scan is segmented: a lane holding a segment start stops the carry);
//   * H_fp = max(non-insertion, I_fp) gives bits 0-4 exactly as the first pass does; the lazy-F
//     pass raises a cell iff I_ex > H_fp (bit 9 for int16; for int32 the word is overwritten
//     with 1 | ext, ext taken from the segment that delivered the carry), and H = max(., I_ex).
// A work-group of NW <= 16 waves solves one problem (NW*64 lanes, C <= 4 columns per lane); waves
// exchange their scan totals and boundary columns through LDS with two barriers per row.
// Values live in 32-bit registers with the reference's arithmetic (int16 saturating or int32
// wrapping). Trace words go to HBM, one uint16 per real cell; the backtrack walks them on the host.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>
#include "nsh_internal.h"
#include "prep.h"

namespace {

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

constexpr int CMAXG = 4, NWMAX = 16;      // al <= 64 * 4 * 16 = 4096

struct GlobProb {
	int64_t rec_off, aa_off, tr_off, pssm_off;
	int32_t nl, al, io, out;
	int32_t C, NW, slen, left;
};
struct GlobArgs {
	const GlobProb *prob;
	const uint32_t *rec; const uint8_t *aas; const int8_t *pssm; const int8_t *mat; int32_t asize;
	int32_t go, ge, fs;
	uint16_t *trace;        // per problem: nl rows x al columns (rows 0, 1 unused)
	int32_t *score;         // per problem: H(nl-1, al-1) (nasw-sse.c:559/962)
};

// LDS exchange area of a work-group (besides the profile)
struct Xch {
	int nbr[2][NWMAX];            // H of the wave's last column, by row parity
	int lfp[2][NWMAX][2];         // I_fp, H_fp of the wave's last column, by row parity
	int tex[NWMAX], tfp[NWMAX], ffp[NWMAX];
	int seg[8];                   // int32: O and flag of segments 0..2
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

template <class AR, int C, bool TRACE>
DEV void glob_solve(const GlobArgs &a, const GlobProb &P, int8_t *prof, Xch &X)
{
	const int NEG = AR::NEG;
	const int L = threadIdx.x & 63, w = threadIdx.x >> 6, NW = P.NW, G = threadIdx.x;
	const int nl = P.nl, al = P.al, slen = P.slen, asz = a.asize;
	const int go = a.go, ge = a.ge, fs = a.fs, io = P.io, goe = go + ge;
	const uint8_t *aas = a.aas + P.aa_off;
	const int8_t *pssm = P.pssm_off >= 0? a.pssm + P.pssm_off : nullptr;
	const uint32_t *rec = a.rec + P.rec_off;
	const int W = 64 * C * NW;                   // columns covered by the work-group
	// column constants
	int jpos[C];
	bool valid[C];
#pragma unroll
	for (int k = 0; k < C; ++k) { const int c = G * C + k; valid[k] = c < al; jpos[k] = c % slen; }
	int lane_reset = 0;
#pragma unroll
	for (int k = 0; k < C; ++k) lane_reset |= valid[k] && jpos[k] == 0;
	// int8 profile in LDS: prof[r * W + c]
#pragma unroll
	for (int k = 0; k < C; ++k) {
		const int c = G * C + k;
		const int aa = valid[k]? aas[c] : 0;
		for (int r = 0; r < 22; ++r)
			prof[r * W + c] = valid[k]? (pssm? pssm[(P.left? al - 1 - c : c) * asz + r] : a.mat[r * asz + aa]) : 0;
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

	auto row = [&](auto Rc) {
		constexpr int R = decltype(Rc)::value, R2 = (R + 1) % 3, R1 = (R + 2) % 3;
		int *H3 = Hs[R], *H2 = Hs[R2], *H1 = Hs[R1], *D3 = Ds[R];
		__syncthreads();   // B1: last row's boundary values are in LDS
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
		const int n1 = nh[R1], n2 = nh[R2], n3 = nh[R];

		// phase A
		int Mv[C], rest[C], yrest[C], bits[C], tgo[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int h3p = k? H3[k - 1] : n3, h1p = k? H1[k - 1] : n1, h2p = k? H2[k - 1] : n2;
			int zz = 0, t, u;
			Mv[k] = AR::add(h3p, (int)pr[k]);
			u = AR::sub(H3[k], go);
			if (D3[k] > u) zz |= 1 << 5;
			const int dn = AR::sub(max(u, D3[k]), gei);
			t = AR::sub(AR::sub(H1[k], io), dm1);
			if (A[k] > t) zz |= 1 << 6;
			A[k] = max(t, A[k]);
			const int av = AR::sub(A[k], a0);
			u = AR::sub(h1p, io);
			t = AR::sub(u, d0);
			if (B[k] > t) zz |= 1 << 7;
			B[k] = max(t, B[k]);
			const int bv = AR::sub(B[k], am2);
			t = AR::sub(u, dp1);
			if (Cc[k] > t) zz |= 1 << 8;
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
			Ifp = (valid[k] && jpos[k] == 0)? fp_start : AR::sub(max(tgo[k - 1], Ifp), ge);
		}
		int Sex = scan_maxplus<AR>(AR::sub(max(tgo[C - 1], Iex), ge), geC, L);
		int Fl = lane_reset;
		int Sfp = scan_maxplus_seg<AR>(AR::sub(max(tgo[C - 1], Ifp), ge), Fl, geC, L);
		if (L == 63) X.tex[w] = Sex, X.tfp[w] = Sfp, X.ffp[w] = Fl;
		__syncthreads();   // B2: wave totals
		// carries entering this wave (the column before it; column -1 for wave 0)
		int cw_ex = fp_start, cw_fp = fp_start;
		for (int v = 0; v < w; ++v) {
			cw_ex = max(X.tex[v], AR::sub(cw_ex, geW));
			cw_fp = X.ffp[v]? X.tfp[v] : max(X.tfp[v], AR::sub(cw_fp, geW));
		}
		int cin_ex = dppk<WAVE_SHR1>(Sex), cin_fp = dppk<WAVE_SHR1>(Sfp), fx = dppk<WAVE_SHR1>(Fl);
		if (L == 0) cin_ex = cw_ex, cin_fp = cw_fp;
		else {
			cin_ex = max(cin_ex, AR::sub(cw_ex, geC * L));
			if (!fx) cin_fp = max(cin_fp, AR::sub(cw_fp, geC * L));
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
				ifp = AR::sub(max(tgo[k - 1], ifp), ge);
			}
			const bool reset = valid[k] && jpos[k] == 0;
			if (reset) ifp = fp_start;
			int zz = bits[k];
			if (reset) { if (AR::IS32) zz |= 1 << 4; }
			else if (k && ifp_prev > AR::sub(hfp_prev, go)) zz |= 1 << 4;
			const int m = Mv[k], r = rest[k];
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
		// bit 4 of slot 0: previous lane's last column (inside the wave)
		{
			const int ipl = dppk<WAVE_SHR1>(ifp_prev), hpl = dppk<WAVE_SHR1>(hfp_prev);
			const bool reset0 = valid[0] && jpos[0] == 0;
			if (L > 0 && !(AR::IS32 && fires0) && !reset0 && ipl > AR::sub(hpl, go)) z[0] |= 1 << 4;
		}
		if (TRACE) {
			uint16_t *dst = tr + (int64_t)i * al + G * C;
#pragma unroll
			for (int k = 0; k < C; ++k)
				if (valid[k] && (k || L || w == 0)) dst[k] = (uint16_t)z[k];
			if (L == 0 && w > 0) z0_pend = z[0], f0_pend = fires0, r0_pend = i;
		}
		if (L == 63) {
			X.nbr[i & 1][w] = Hf[C - 1];
			X.lfp[i & 1][w][0] = ifp_prev, X.lfp[i & 1][w][1] = hfp_prev;
		}
#pragma unroll
		for (int k = 0; k < C; ++k) H3[k] = Hf[k];
		const int Vs = dppk<WAVE_SHR1>(Hf[C - 1]);
		nh[R] = L == 0? NEG : Vs;                 // lane 0 of waves > 0 is filled from LDS next row
		if (i == 2 && G == 0) nh[R2] = NEG, nh[R1] = NEG;
		++i;
	};
	if (nl > 2) {
		for (;;) {
			row(std::integral_constant<int, 0>()); if (i >= nl) break;
			row(std::integral_constant<int, 1>()); if (i >= nl) break;
			row(std::integral_constant<int, 2>()); if (i >= nl) break;
		}
		__syncthreads();
		if (TRACE && L == 0 && w > 0) {
			const int ipl = X.lfp[(i - 1) & 1][w - 1][0], hpl = X.lfp[(i - 1) & 1][w - 1][1];
			const bool reset0 = valid[0] && jpos[0] == 0;
			if (!(AR::IS32 && f0_pend) && !reset0 && ipl > AR::sub(hpl, go)) z0_pend |= 1 << 4;
			if (valid[0]) tr[(int64_t)r0_pend * al + G * C] = (uint16_t)z0_pend;
		}
	}
	// score = H(nl-1, al-1); the last row computed sits in Hs[(nl-3) % 3]
	const int last = nl > 2? (nl - 3) % 3 : -1;
	const int ge_ = (al - 1) / C, ke = (al - 1) % C;
	if (G == ge_) {
		int sc = NEG;
#pragma unroll
		for (int k = 0; k < C; ++k)
			if (k == ke) sc = last == 0? Hs[0][k] : last == 1? Hs[1][k] : last == 2? Hs[2][k] : NEG;
		a.score[P.out] = sc;
	}
}

// MAXT: the work-group size bound (512 keeps the largest column count free of spills; 1024 serves
// proteins longer than 2048)
template <class AR, bool TRACE, int MAXT>
__global__ void __launch_bounds__(MAXT) glob_kernel(GlobArgs a)
{
	extern __shared__ int8_t prof[];
	__shared__ Xch X;
	const GlobProb P = a.prob[blockIdx.x];
	switch (P.C) {
	case 1: glob_solve<AR, 1, TRACE>(a, P, prof, X); break;
	case 2: glob_solve<AR, 2, TRACE>(a, P, prof, X); break;
	case 3: glob_solve<AR, 3, TRACE>(a, P, prof, X); break;
	default: glob_solve<AR, 4, TRACE>(a, P, prof, X); break;
	}
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
#define HIPOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("%s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

// nasw-sse.c:34-42
void fix_tiny_uv(int32_t n, uint32_t *cigar)
{
	for (int32_t k = 0; k < n; ++k) {
		const int32_t op = cigar[k] & 0xf;
		if ((op == NS_CIGAR_U || op == NS_CIGAR_V) && cigar[k] >> 4 < 3) cigar[k] = cigar[k] >> 4 << 4 | NS_CIGAR_G;
	}
}

// walk the trace words from (nl-1, al-1) (nasw-sse.c:44-93); tr[i*al + j]
void backtrack(void *km, const uint16_t *tr, int32_t nl, int32_t al, ns_rst_t *r)
{
	int32_t i = nl - 1, j = al - 1, last = 0;
	uint32_t *cigar = r->cigar;
	while (i >= 2 && j >= 0) {
		int32_t x = tr[(int64_t)i * al + j], state, ext;
		if (x >> 9 & 1) x = 1 | x >> 4 << 4;
		state = last == 0? x & 0xf : last;
		ext = state >= 1 && state <= 5? x >> (state + 3) & 1 : 0;
		switch (state) {
		case 0: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_M, 1), i -= 3, --j; break;
		case 1: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_I, 1), --j; break;
		case 2: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_D, 1), i -= 3; break;
		case 3: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_N, 1), --i; break;
		case 4: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_U, 1), --i; if (!ext) --j; break;
		case 5: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_V, 1), --i; if (!ext) --j; break;
		case 6: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_F, 1), --i; break;
		case 7: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_F, 2), i -= 2; break;
		case 8: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_G, 1), --i, --j; break;
		case 9: cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_G, 2), i -= 2, --j; break;
		default: i = -1; break;   // not produced by the kernel
		}
		last = state >= 1 && state <= 5 && ext? state : 0;
	}
	if (j > 0) cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_I, j);
	if (i >= 0) {
		const int32_t l = (i + 1) / 3, t = (i + 1) % 3;
		if (l > 0) cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_D, l);
		if (t != 0) cigar = ns_push_cigar(km, &r->n_cigar, &r->m_cigar, cigar, NS_CIGAR_F, t);
	}
	for (int32_t k = 0; k < r->n_cigar >> 1; ++k) std::swap(cigar[k], cigar[r->n_cigar - 1 - k]);
	fix_tiny_uv(r->n_cigar, cigar);
	r->cigar = cigar;
}

template <typename T>
hipError_t dalloc(T **p, size_t n) { return hipMalloc((void**)p, n * sizeof(T) + 16); }

std::mutex g_glob_mu;   // one problem at a time through the shared stream

} // namespace

int nsh_global_run(void *km, const char *ns, int32_t nl, const char *as, int32_t al,
		const ns_opt_t *opt, const uint8_t *ss, int32_t is32, ns_rst_t *r)
{
	if (!nsh_ensure_device()) return 0;
	const bool cigar = (opt->flag & NS_F_CIGAR) && !(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	if (!is32 && !cigar) { nsh_set_error("score-only int16 DP goes through the extension engine"); return 0; }
	if (al < 1 || nl < 0) { nsh_set_error("global DP: nl=%d al=%d", nl, al); return 0; }
	if (opt->asize != 22) { nsh_set_error("asize %d not supported", opt->asize); return 0; }
	const int32_t NW = (al + 64 * CMAXG - 1) / (64 * CMAXG);
	if (NW > NWMAX) { nsh_set_error("global DP: protein length %d exceeds %d", al, 64 * CMAXG * NWMAX); return 0; }
	const int32_t C = (al + 64 * NW - 1) / (64 * NW);
	if (opt->ge < 0 || opt->go < 0 || opt->ge > 32767 || opt->go > 32767 || opt->fs < 0 || opt->fs > 32767) {
		nsh_set_error("gap penalties go=%d ge=%d fs=%d out of the supported range", opt->go, opt->ge, opt->fs); return 0;
	}
	for (int t = 0; t < 64; ++t)
		if (opt->codon[t] >= 22) { nsh_set_error("codon table entry %d out of range", t); return 0; }
	for (int t = 0; t < 256; ++t)
		if (opt->aa20[t] >= 22 || opt->nt4[t] > 4) { nsh_set_error("alphabet table entry %d out of range", t); return 0; }
	const int vs = is32? 4 : 8;
	std::lock_guard<std::mutex> lk(g_glob_mu);
	static hipStream_t stream = nullptr;
	uint8_t *d_ns = nullptr, *d_ss = nullptr, *d_as = nullptr, *d_aas = nullptr;
	uint32_t *d_rec = nullptr;
	int8_t *d_mat = nullptr, *d_pssm = nullptr;
	uint16_t *d_tr = nullptr;
	int32_t *d_score = nullptr;
	GlobProb *d_prob = nullptr;
	nsh_prep::PrepProb *d_pp = nullptr;
	nsh_prep::PrepConst *d_pc = nullptr;
	int ok = 0;
	std::vector<uint16_t> tr;
	const auto t0 = std::chrono::steady_clock::now();
	{
		if (!stream) {
			HIPOK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
			const int big = 22 * 64 * CMAXG * NWMAX;   // profile bytes of the largest work-group
			HIPOK(hipFuncSetAttribute(reinterpret_cast<const void*>(&glob_kernel<Arith16, true, 1024>), hipFuncAttributeMaxDynamicSharedMemorySize, big));
			HIPOK(hipFuncSetAttribute(reinterpret_cast<const void*>(&glob_kernel<Arith32, true, 1024>), hipFuncAttributeMaxDynamicSharedMemorySize, big));
			HIPOK(hipFuncSetAttribute(reinterpret_cast<const void*>(&glob_kernel<Arith32, false, 1024>), hipFuncAttributeMaxDynamicSharedMemorySize, big));
		}
		nsh_prep::PrepProb pp = {};
		pp.ns_off = 0, pp.as_off = 0, pp.rec_off = 0, pp.aa_off = 0;
		pp.nl = nl, pp.al = al, pp.left = (opt->flag & NS_F_EXT_LEFT) != 0, pp.io = opt->io, pp.has_ss = ss != nullptr;
		nsh_prep::PrepConst pc;
		std::memcpy(pc.nt4, opt->nt4, 256), std::memcpy(pc.aa20, opt->aa20, 256), std::memcpy(pc.codon, opt->codon, 64);
		for (int t = 0; t < 6; ++t) pc.sp[t] = opt->sp[t];
		pc.sp_null = opt->sp_null_bonus;
		GlobProb gp = {};
		gp.rec_off = 0, gp.aa_off = 0, gp.tr_off = 0, gp.pssm_off = opt->pssm? 0 : -1;
		gp.nl = nl, gp.al = al, gp.io = opt->io, gp.out = 0, gp.C = C, gp.NW = NW, gp.slen = (al + vs - 1) / vs, gp.left = pp.left;
		HIPOK(dalloc(&d_ns, nl + 1)); HIPOK(dalloc(&d_as, al + 1)); HIPOK(dalloc(&d_aas, al + 1)); HIPOK(dalloc(&d_rec, nl + 1));
		if (ss) HIPOK(dalloc(&d_ss, nl + 1));
		HIPOK(dalloc(&d_mat, 22 * 22));
		if (opt->pssm) HIPOK(dalloc(&d_pssm, (size_t)al * 22));
		if (cigar) HIPOK(dalloc(&d_tr, (size_t)nl * al));
		HIPOK(dalloc(&d_score, 1)); HIPOK(dalloc(&d_prob, 1)); HIPOK(dalloc(&d_pp, 1)); HIPOK(dalloc(&d_pc, 1));
		if (nl) HIPOK(hipMemcpyAsync(d_ns, ns, nl, hipMemcpyHostToDevice, stream));
		HIPOK(hipMemcpyAsync(d_as, as, al, hipMemcpyHostToDevice, stream));
		if (ss && nl) HIPOK(hipMemcpyAsync(d_ss, ss, nl, hipMemcpyHostToDevice, stream));
		HIPOK(hipMemcpyAsync(d_mat, opt->sc, 22 * 22, hipMemcpyHostToDevice, stream));
		if (opt->pssm) HIPOK(hipMemcpyAsync(d_pssm, opt->pssm, (size_t)al * 22, hipMemcpyHostToDevice, stream));
		HIPOK(hipMemcpyAsync(d_prob, &gp, sizeof(gp), hipMemcpyHostToDevice, stream));
		HIPOK(hipMemcpyAsync(d_pp, &pp, sizeof(pp), hipMemcpyHostToDevice, stream));
		HIPOK(hipMemcpyAsync(d_pc, &pc, sizeof(pc), hipMemcpyHostToDevice, stream));
		nsh_prep::launch(d_pp, 1, d_ns, d_ss, d_as, d_rec, d_aas, d_pc, stream);
		GlobArgs ga;
		ga.prob = d_prob, ga.rec = d_rec, ga.aas = d_aas, ga.pssm = d_pssm, ga.mat = d_mat, ga.asize = 22;
		ga.go = opt->go, ga.ge = opt->ge, ga.fs = opt->fs, ga.trace = d_tr, ga.score = d_score;
		const size_t lds = (size_t)22 * 64 * C * NW;
		const dim3 grid(1), block(64 * NW);
		if (NW <= 8) {
			if (!is32) hipLaunchKernelGGL((glob_kernel<Arith16, true, 512>), grid, block, lds, stream, ga);
			else if (cigar) hipLaunchKernelGGL((glob_kernel<Arith32, true, 512>), grid, block, lds, stream, ga);
			else hipLaunchKernelGGL((glob_kernel<Arith32, false, 512>), grid, block, lds, stream, ga);
		} else {
			if (!is32) hipLaunchKernelGGL((glob_kernel<Arith16, true, 1024>), grid, block, lds, stream, ga);
			else if (cigar) hipLaunchKernelGGL((glob_kernel<Arith32, true, 1024>), grid, block, lds, stream, ga);
			else hipLaunchKernelGGL((glob_kernel<Arith32, false, 1024>), grid, block, lds, stream, ga);
		}
		HIPOK(hipGetLastError());
		nsh_count_launch(2);
		int32_t score = 0;
		HIPOK(hipMemcpyAsync(&score, d_score, 4, hipMemcpyDeviceToHost, stream));
		if (cigar) {
			tr.resize((size_t)nl * al);
			HIPOK(hipMemcpyAsync(tr.data(), d_tr, tr.size() * 2, hipMemcpyDeviceToHost, stream));
		}
		HIPOK(hipStreamSynchronize(stream));
		r->n_cigar = 0, r->nt_len = nl, r->aa_len = al, r->score = nl > 2? score : (is32? -0x40000000 : -32768);
		if (cigar) backtrack(km, tr.data(), nl, al, r);
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_ns, d_ss, d_as, d_aas, d_rec, d_mat, d_pssm, d_tr, d_score, d_prob, d_pp, d_pc };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}
