// Device side of the batched extension DP (see ext_dp.hip for the execution model).
// Included by ext_dp_k8.hip and ext_dp_k16.hip, which instantiate the two LDS variants.
#ifndef NSH_EXT_DP_IMPL_H
#define NSH_EXT_DP_IMPL_H
#include <hip/hip_runtime.h>
#include <climits>
#include <type_traits>
#include "nsh_internal.h"
#include "ext_dp_types.h"

namespace nsh_ext {

typedef short s2 __attribute__((ext_vector_type(2)));
#define DEV static __device__ __forceinline__

DEV s2 adds(s2 a, s2 b) { return __builtin_elementwise_add_sat(a, b); }
DEV s2 subs(s2 a, s2 b) { return __builtin_elementwise_sub_sat(a, b); }
DEV s2 vmax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
DEV s2 vmin(s2 a, s2 b) { return __builtin_elementwise_min(a, b); }
DEV int pk(s2 a) { return __builtin_bit_cast(int, a); }
DEV s2 up(int a) { return __builtin_bit_cast(s2, a); }
DEV s2 spl(int x) { return s2{(short)x, (short)x}; }
DEV int lo16(int x) { return (int)(short)(x & 0xffff); }
DEV int hi16(int x) { return (int)(short)((unsigned)x >> 16); }

constexpr int NEG = -32768;
constexpr int PACKNEG = (int)0x80008000u;
constexpr int CMAX_SMALL = 8;    // columns per lane per half: npad <= 1024 (22.5 KB LDS profile)
constexpr int CMAX = 16;         // large variant: npad <= 2048 (45 KB LDS profile)
constexpr int NRES = 22;         // residue codes of the profile rows

// DPP controls (GFX9 encoding)
enum : int { ROW_SHR1 = 0x111, ROW_SHR2 = 0x112, ROW_SHR4 = 0x114, ROW_SHR8 = 0x118,
	WAVE_SHR1 = 0x138, ROW_BCAST15 = 0x142, ROW_BCAST31 = 0x143 };
// lanes without a source keep their own value (old == src, bound_ctrl off)
template <int CTRL, int ROWMASK = 0xf>
DEV int dpp(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, ROWMASK, 0xf, false); }

template <int C>
DEV void ext_solve(const ExtArgs &a, const ExtProb &P, int8_t *prof, const int L)
{
	constexpr int HALF = 64 * C, ROWB = 128 * C;
	const int nl = P.nl, al = P.al, npad = P.npad, asz = a.asize;
	const int go = a.go, ge = a.ge, fs = a.fs, io = P.io, eb = a.eb, xdrop = a.xdrop;
	const uint8_t *aas = a.aas + P.aa_off;
	const int8_t *pssm = P.pssm_off >= 0? a.pssm + P.pssm_off : nullptr;
	const uint32_t *rec = a.rec + P.rec_off;
	const int pen_len = 3 * al;

	// ---- per-lane column constants and the int8 query profile in LDS ----
	// sel[k]: v_perm selector packing the two sign-extended profile bytes of column k into an
	// int16 pair; padded columns (al <= c < npad) become -32768 like the reference profile.
	// lim[k]: columns at or beyond npad (not part of the reference's row maximum) are masked.
	uint32_t sel[C];
	s2 lim[C];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		const int clo = L * C + k, chi = HALF + L * C + k;
		const bool plo = clo >= al && clo < npad, phi = chi >= al && chi < npad;
		sel[k] = (plo? 0x000Cu : 0x0100u) | (phi? 0x040C0000u : 0x05040000u);
		lim[k] = s2{(short)(clo < npad? 32767 : NEG), (short)(chi < npad? 32767 : NEG)};
	}
	__syncthreads();
#pragma unroll
	for (int h = 0; h < 2; ++h)
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int c = h * HALF + L * C + k;
			const bool real = c < al;
			const int q = P.left? al - 1 - c : c;
			const int aa = real? aas[c] : 0;
			int8_t *dst = prof + h * HALF + k * 64 + L;
			for (int r = 0; r < NRES; ++r) {
				int v = -128;
				if (real) v = pssm? pssm[q * asz + r] : a.mat[r * asz + aa];
				dst[r * ROWB] = (int8_t)v;
			}
		}
	__syncthreads();

	// ---- scan constants: the insertion carry decays by ge per column, geC per lane ----
	const int geC = ge * C;
	const s2 DEC1 = spl(geC), DEC2 = spl(2 * geC), DEC4 = spl(4 * geC), DEC8 = spl(8 * geC);
	const s2 DEC15 = spl(geC * ((L & 15) + 1));
	const s2 DEC31 = spl(L >= 32? geC * (L - 31) : 0);
	const s2 DECL = s2{0, (short)(geC * L)};
	const s2 GO = spl(go), GE = spl(ge), IO = spl(io), FS = spl(fs);

	// ---- DP state: three rows of H and D, intron states, best-row snapshot ----
	s2 Hs[3][C], Ds[3][C], A[C], B[C], Cc[C], Hmax[C], nh[3];
#pragma unroll
	for (int k = 0; k < C; ++k) {
		Hs[0][k] = Hs[1][k] = Hs[2][k] = Ds[0][k] = Ds[1][k] = Ds[2][k] = up(PACKNEG);
		A[k] = B[k] = Cc[k] = Hmax[k] = up(PACKNEG);
	}
	// left neighbour column of each lane's first column, rows i-3 (nh[0]), i-2 (nh[1]), i-1 (nh[2])
	// at the first row; column -1 holds H(-1,-1)=0 and H(0,-1)=H(1,-1)=-fs only for row 2.
	nh[0] = nh[1] = nh[2] = up(PACKNEG);
	if (L == 0) nh[0] = s2{0, (short)NEG}, nh[1] = nh[2] = s2{(short)-fs, (short)NEG};

	// end column (al-1): half, lane and slot
	const int he = (al - 1) >= HALF, qe = (al - 1) - he * HALF, Le = qe / C, ke = qe % C;

	// ---- row records and penalties, 64 rows per VGPR, double buffered ----
	auto load_blk = [&](int i0, int &rv, int &pv) {
		const int i = i0 + L;
		rv = i < nl? (int)rec[i] : 0;
		const int d = i - pen_len;
		pv = d < 2? 0 : a.pen[d < a.pen_n? d : a.pen_n - 1];
	};
	int recv, penv, recn = 0, penn = 0;
	load_blk(0, recv, penv);
	if (nl > 64) load_blk(64, recn, penn);
	int d0 = 0, dp1 = 0, am1 = 0, a0 = 0, dm1 = 0, am2 = 0;
	if (nl >= 2) {
		const int r0 = __builtin_amdgcn_readlane(recv, 0), r1 = __builtin_amdgcn_readlane(recv, 1);
		d0 = (int)(int8_t)(r0 >> 8), dp1 = (int)(int8_t)(r1 >> 8);
		am1 = (int)(int8_t)(r0 >> 16), a0 = (int)(int8_t)(r1 >> 16);
	}

	int best = INT_MIN, best_log = INT_MIN, best_i = -1, last_i = 1, i = 2;

	auto row = [&](auto Rc) -> bool {
		constexpr int R = decltype(Rc)::value, R2 = (R + 1) % 3, R1 = (R + 2) % 3;
		s2 *H3 = Hs[R], *H2 = Hs[R2], *H1 = Hs[R1], *D3 = Ds[R];
		if ((i & 63) == 0) { // next block of records
			recv = recn, penv = penn;
			if (i + 64 < nl) load_blk(i + 64, recn, penn);
		}
		const int rc = __builtin_amdgcn_readlane(recv, i & 63);
		const int pen_i = __builtin_amdgcn_readlane(penv, i & 63);
		const int res = rc & 0xff;
		dm1 = d0, d0 = dp1, dp1 = (int)(int8_t)(rc >> 8);
		am2 = am1, am1 = a0, a0 = (int)(int8_t)(rc >> 16);
		const s2 GEI = spl(res == 20? fs : ge);
		const s2 DM1 = spl(dm1), D0 = spl(d0), DP1 = spl(dp1), A0 = spl(a0), AM1 = spl(am1), AM2 = spl(am2);

		// profile of this row's residue
		const int8_t *pr = prof + res * ROWB + L;
		s2 s[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int lo = pr[k * 64], hi = pr[HALF + k * 64];
			s[k] = up((int)__builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, sel[k]));
		}

		// local pass: all terms, insertion chain with zero carry
		const s2 n1 = nh[R1], n2 = nh[R2], n3 = nh[R];
		s2 uprev = subs(n1, IO), mprev = vmax(n1, n2), h3prev = n3;
		s2 hprev = up(PACKNEG), I = up(PACKNEG), rm = up(PACKNEG);
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const s2 h3 = H3[k], h1 = H1[k];
			s2 h = adds(h3prev, s[k]);                                   // H(i-3,j-1) + s
			const s2 dn = subs(vmax(subs(h3, GO), D3[k]), GEI);           // D
			h = vmax(h, dn);
			const s2 u = subs(h1, IO);
			A[k] = vmax(subs(u, DM1), A[k]);     h = vmax(h, subs(A[k], A0));   // phase-0 intron
			B[k] = vmax(subs(uprev, D0), B[k]);  h = vmax(h, subs(B[k], AM2));  // phase-1
			Cc[k] = vmax(subs(uprev, DP1), Cc[k]); h = vmax(h, subs(Cc[k], AM1)); // phase-2
			const s2 m = vmax(h1, H2[k]);
			h = vmax(h, subs(vmax(m, mprev), FS));                          // 4 frameshift moves
			I = subs(vmax(subs(hprev, GO), I), GE);                          // insertion (local)
			h = vmax(h, I);
			rm = vmax(rm, vmin(h, lim[k]));
			H3[k] = h, D3[k] = dn;
			hprev = h, uprev = u, mprev = m, h3prev = h3;
		}

		// max-plus scan of the lanes' outgoing insertions
		int S = pk(subs(vmax(subs(hprev, GO), I), GE));
		S = pk(vmax(up(S), subs(up(dpp<ROW_SHR1>(S)), DEC1)));
		S = pk(vmax(up(S), subs(up(dpp<ROW_SHR2>(S)), DEC2)));
		S = pk(vmax(up(S), subs(up(dpp<ROW_SHR4>(S)), DEC4)));
		S = pk(vmax(up(S), subs(up(dpp<ROW_SHR8>(S)), DEC8)));
		S = pk(vmax(up(S), subs(up(dpp<ROW_BCAST15, 0xA>(S)), DEC15)));
		S = pk(vmax(up(S), subs(up(dpp<ROW_BCAST31, 0xC>(S)), DEC31)));
		const int Tlo = lo16(__builtin_amdgcn_readlane(S, 63));   // insertion leaving the low half
		int Sx = dpp<WAVE_SHR1>(S);
		if (L == 0) Sx = PACKNEG;
		const s2 cin = vmax(up(Sx), subs(s2{(short)NEG, (short)Tlo}, DECL));

		// fold the carry into the lane's cells
		s2 c = cin;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			H3[k] = vmax(H3[k], c);
			if (k + 1 < C) c = subs(c, GE);
		}
		rm = vmax(rm, vmin(cin, lim[0]));

		// row maximum over both halves
		int m = pk(rm);
		m = pk(vmax(up(m), up(dpp<ROW_SHR1>(m))));
		m = pk(vmax(up(m), up(dpp<ROW_SHR2>(m))));
		m = pk(vmax(up(m), up(dpp<ROW_SHR4>(m))));
		m = pk(vmax(up(m), up(dpp<ROW_SHR8>(m))));
		m = pk(vmax(up(m), up(dpp<ROW_BCAST15, 0xA>(m))));
		m = pk(vmax(up(m), up(dpp<ROW_BCAST31, 0xC>(m))));
		const int mt = __builtin_amdgcn_readlane(m, 63);
		const int rowmax = std::max(lo16(mt), hi16(mt));

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

		// row score, length penalty, best row and x-drop (nasw-sse.c:438-447)
		const int es = ev + eb;
		const int tmp = rowmax > es? rowmax : es;
		const int tlog = tmp - pen_i;
		if (tlog > best_log) {
			best = tmp, best_log = tlog, best_i = i;
#pragma unroll
			for (int k = 0; k < C; ++k) Hmax[k] = H3[k];
		}
		// left neighbour of the next row: lane L-1's last column; lane 0 gets column -1
		// (neg_inf for rows > 2) in the low half and the low half's last column in the high half
		const int V = pk(H3[C - 1]);
		const int Vs = dpp<WAVE_SHR1>(V);
		const int V63 = __builtin_amdgcn_readlane(V, 63);
		nh[R] = up(L == 0? (int)(((unsigned)V63 << 16) | 0x8000u) : Vs);
		if (i == 2 && L == 0) nh[R2].x = NEG, nh[R1].x = NEG;
		(void)best;
		last_i = i;
		if (best_log - tlog > xdrop) return false;
		return ++i < nl;
	};

	if (nl > 2) {
		for (;;) {
			if (!row(std::integral_constant<int, 0>())) break;
			if (!row(std::integral_constant<int, 1>())) break;
			if (!row(std::integral_constant<int, 2>())) break;
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
				const int v = (h? (int)Hmax[k].y : (int)Hmax[k].x) + (j == al - 1? eb : 0);
				if (v > bv) bv = v, bj = j;
			}
		}
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) {
		const int ov = __shfl_xor(bv, off), oj = __shfl_xor(bj, off);
		if (ov > bv || (ov == bv && oj < bj)) bv = ov, bj = oj;
	}
	if (L == 0) {
		int32_t *o = a.out + 4 * (int64_t)P.out;
		o[0] = best_i + 1, o[1] = bj + 1, o[2] = bv, o[3] = last_i - 1;
	}
}

template <int CM>
__global__ void __launch_bounds__(64) ext_dp_kernel(ExtArgs a)
{
	__shared__ int8_t prof[NRES * 128 * CM];
	const int L = threadIdx.x;
	for (;;) {
		int p = 0;
		if (L == 0) p = atomicAdd(a.counter, 1);
		p = __builtin_amdgcn_readlane(p, 0);
		if (p >= a.n) break;
		const ExtProb P = a.prob[p];
		if (CM <= CMAX_SMALL) {
			switch (P.C) {
			case 1: ext_solve<1>(a, P, prof, L); break;
			case 2: ext_solve<2>(a, P, prof, L); break;
			case 3: ext_solve<3>(a, P, prof, L); break;
			case 4: ext_solve<4>(a, P, prof, L); break;
			case 5: ext_solve<5>(a, P, prof, L); break;
			case 6: ext_solve<6>(a, P, prof, L); break;
			case 7: ext_solve<7>(a, P, prof, L); break;
			default: ext_solve<CMAX_SMALL>(a, P, prof, L); break;
			}
		} else {
			switch (P.C) {
			case 9: ext_solve<9>(a, P, prof, L); break;
			case 10: ext_solve<10>(a, P, prof, L); break;
			case 11: ext_solve<11>(a, P, prof, L); break;
			case 12: ext_solve<12>(a, P, prof, L); break;
			case 13: ext_solve<13>(a, P, prof, L); break;
			case 14: ext_solve<14>(a, P, prof, L); break;
			case 15: ext_solve<15>(a, P, prof, L); break;
			default: ext_solve<CMAX>(a, P, prof, L); break;
			}
		}
	}
}

} // namespace nsh_ext
#endif
