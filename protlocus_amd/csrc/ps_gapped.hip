// Banded local affine alignment of a profile against ORFs: ps_gapped (prescreen.c:346-382),
// the gapped stage of the prescreen search, for many (ORF, diagonal) items in one launch.
//
// One wave per item. The wave holds one row of the band: 64·C columns, C consecutive per lane.
// The band moves one column right per row, so it is kept in band-relative coordinates: a cell's
// diagonal predecessor is the same slot of the previous row, and its vertical predecessor is the
// next slot (a lane shift). A column the band has not reached yet enters as H = 0, F = -inf, as in
// the reference's arrays. For band <= 0 the full row is held in ORF coordinates instead.
//
// The horizontal gap chain E(t+1) = max(E(t) - ge, h(t) - go) needs go >= ge, so that a cell's
// own E never opens a new gap. Each lane runs it over its columns from -inf; a max-plus scan
// across lanes (decay ge·C per lane) then gives each lane its exact incoming E. Int32 scores, -inf
// = -0x3fffffff, as the reference.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr int NEG = -0x3fffffff;
constexpr int CMAX = 16;

// a DPP lane move: lanes whose source is outside the pattern (or whose row is masked off) keep old
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp(int old, int src)
{
	return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWS, 0xf, false);
}
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;        // lane i takes lane i + 1 / i - 1
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;    // lane 15 / 31 of the row below, to rows 1,3 / 2,3

struct GapArgs {
	const int8_t *pssm; int32_t L;
	const int64_t *prow; const int32_t *plen;   // per item: first PSSM row and length (null: pssm, L for all)
	const uint8_t *aa; const int64_t *off; const int32_t *len, *diag;
	int32_t band, go, ge;
	const int32_t *items; int32_t n_items;   // item indices of this launch (one column count)
	int32_t *out;
};

template <int C, bool FULL>
__global__ void __launch_bounds__(256) ps_gapped_kernel(GapArgs a)
{
	__shared__ int8_t stage[4][64 * 22];   // per wave: 64 consecutive PSSM rows
	const int Ln = threadIdx.x & 63;
	const int wv = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int w = blockIdx.x * 4 + wv;
	if (w >= a.n_items) return;   // wave-uniform
	const int item = a.items[w];
	const uint8_t *orf = a.aa + a.off[item];
	const int n = a.len[item], dg = a.diag[item], band = a.band, go = a.go, ge = a.ge;
	const int8_t *pssm = a.prow? a.pssm + a.prow[item] * 22 : a.pssm;
	const int L = a.prow? a.plen[item] : a.L;
	const int W = FULL? n : 2 * band + 1;   // columns held
	int8_t *tab = stage[wv];
	int H[C], F[C], res[C];
	// rows whose band meets [0, n): all of them in full mode
	const int i0 = FULL? 0 : std::max(0, -dg - band), i1 = FULL? L : std::min(L, n - dg + band);
	{   // the residue under each slot at row i0 (0 off the ORF); banded slots then move one column per row
		const int base = FULL? 0 : i0 + dg - band;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int t = base + Ln * C + k;
			H[k] = 0, F[k] = NEG, res[k] = t >= 0 && t < n? orf[t] : 0;
		}
	}
	int best = 0, inq = 0;
	for (int i = i0; i < i1; ++i) {
		const int j = (i - i0) & 63;
		if (j == 0) {   // stage rows i..i+63; banded: the residue entering the last slot after each of them
			const int r = i + Ln;
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			if (r < i1) {
				const uint16_t *src = (const uint16_t*)(pssm + (int64_t)r * 22);
				uint16_t *dst = (uint16_t*)(tab + Ln * 22);
#pragma unroll
				for (int b = 0; b < 11; ++b) dst[b] = src[b];
			}
			if (!FULL) {
				const int tn = r + 1 + dg - band + 64 * C - 1;
				inq = tn >= 0 && tn < n? orf[tn] : 0;
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		const int8_t *row = tab + j * 22;
		const int base = FULL? 0 : i + dg - band;   // ORF column of slot 0
		int hup[C], fup[C], hdg[C];
		if (FULL) {   // vertical: same slot; diagonal: the slot to the left (column -1 holds 0)
			const int left = dpp<kWaveShr1>(0, H[C - 1]);
#pragma unroll
			for (int k = 0; k < C; ++k) hup[k] = H[k], fup[k] = F[k], hdg[k] = k? H[k - 1] : left;
		} else {      // vertical: the next slot (beyond the band: H = 0, F = -inf); diagonal: same slot
			const int hn = dpp<kWaveShl1>(0, H[0]), fn = dpp<kWaveShl1>(NEG, F[0]);
#pragma unroll
			for (int k = 0; k < C; ++k) {
				hup[k] = k + 1 < C? H[k + 1] : hn;
				fup[k] = k + 1 < C? F[k + 1] : fn;
				hdg[k] = H[k];
			}
		}
		int hp[C], fn[C], eloc[C];
		bool val[C];
		int e = NEG;
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int u = Ln * C + k, t = base + u;
			val[k] = u < W && t >= 0 && t < n;
			const int s = val[k]? row[res[k]] : 0;
			const int f = std::max(fup[k] - ge, hup[k] - go);
			hp[k] = val[k]? std::max(std::max(hdg[k] + s, f), 0) : NEG;
			fn[k] = val[k]? f : NEG;
			eloc[k] = e;                                  // E entering slot k from this lane's columns
			e = std::max(e - ge, hp[k] - go);
		}
		// incoming E of each lane: max over lanes m < L of the E leaving m, decayed by ge·C per lane
		// (shifts 1-8 inside each row of 16 lanes, then lane 15 and lane 31 broadcast onwards)
		int V = e;
		const int geC = ge * C;
		V = std::max(V, dpp<kRowShr1>(NEG, V) - geC);
		V = std::max(V, dpp<kRowShr2>(NEG, V) - 2 * geC);
		V = std::max(V, dpp<kRowShr4>(NEG, V) - 4 * geC);
		V = std::max(V, dpp<kRowShr8>(NEG, V) - 8 * geC);
		V = std::max(V, dpp<kRowBcast15, 0xa>(NEG, V) - geC * ((Ln & 15) + 1));
		V = std::max(V, dpp<kRowBcast31, 0xc>(NEG, V) - geC * ((Ln & 31) + 1));
		const int inc = dpp<kWaveShr1>(NEG, V);
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int E = std::max(eloc[k], inc - ge * k);
			const int h = val[k]? std::max(hp[k], E) : 0;
			H[k] = h, F[k] = fn[k];
			best = std::max(best, h);
		}
		if (!FULL) {   // next row: every slot takes its right neighbour's residue
			const int nx = dpp<kWaveShl1>(__builtin_amdgcn_readlane(inq, j), res[0]);
#pragma unroll
			for (int k = 0; k + 1 < C; ++k) res[k] = res[k + 1];
			res[C - 1] = nx;
		}
	}
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) best = std::max(best, __shfl_xor(best, d));
	if (Ln == 0) a.out[item] = best;
}

template <int C>
void launch_c(const GapArgs &a, bool full, hipStream_t s)
{
	const dim3 grid((a.n_items + 3) / 4), block(256);
	if (full) hipLaunchKernelGGL((ps_gapped_kernel<C, true>), grid, block, 0, s, a);
	else hipLaunchKernelGGL((ps_gapped_kernel<C, false>), grid, block, 0, s, a);
}

void launch(int C, const GapArgs &a, bool full, hipStream_t s)
{
	switch (C) {
	case 1: launch_c<1>(a, full, s); break;   case 2: launch_c<2>(a, full, s); break;
	case 3: launch_c<3>(a, full, s); break;   case 4: launch_c<4>(a, full, s); break;
	case 5: launch_c<5>(a, full, s); break;   case 6: launch_c<6>(a, full, s); break;
	case 7: launch_c<7>(a, full, s); break;   case 8: launch_c<8>(a, full, s); break;
	case 9: launch_c<9>(a, full, s); break;   case 10: launch_c<10>(a, full, s); break;
	case 11: launch_c<11>(a, full, s); break; case 12: launch_c<12>(a, full, s); break;
	case 13: launch_c<13>(a, full, s); break; case 14: launch_c<14>(a, full, s); break;
	case 15: launch_c<15>(a, full, s); break; default: launch_c<16>(a, full, s); break;
	}
}

std::mutex g_mu;

} // namespace

// device-resident items (the prescreen search): one launch for the items of one column count
void nsh_psg::launch(int C, bool full, const int8_t *d_pssm, int32_t L, const int64_t *d_prow, const int32_t *d_plen,
		const uint8_t *d_aa, const int64_t *d_off, const int32_t *d_len, const int32_t *d_diag, int32_t band, int32_t go,
		int32_t ge, const int32_t *d_items, int32_t n_items, int32_t *d_out, hipStream_t s)
{
	GapArgs ga;
	ga.pssm = d_pssm, ga.L = L, ga.prow = d_prow, ga.plen = d_plen, ga.aa = d_aa, ga.off = d_off, ga.len = d_len, ga.diag = d_diag;
	ga.band = band, ga.go = go, ga.ge = ge, ga.items = d_items, ga.n_items = n_items, ga.out = d_out;
	::launch(C, ga, full, s);
}

#define PGOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("ps_gapped: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" int nsh_ps_gapped_batch(const int8_t *pssm, int32_t L, const uint8_t *aa, int64_t l_aa, int32_t n,
		const int64_t *orf_off, const int32_t *orf_len, const int32_t *diag, int32_t band, int32_t go, int32_t ge,
		int32_t *sw)
{
	if (!nsh_ensure_device()) return 0;
	if (!pssm || L < 0 || !aa || l_aa < 0 || n < 0 || (n > 0 && (!orf_off || !orf_len || !diag || !sw)) || ge < 0 ||
			go < ge || go > 65535 || band > 64 * CMAX / 2 - 1) {
		nsh_set_error("ps_gapped: bad arguments (need 0 <= ge <= go <= 65535, band <= %d)", 64 * CMAX / 2 - 1);
		return 0;
	}
	if (n == 0) return 1;
	const bool full = band <= 0;
	std::vector<std::vector<int32_t>> by_c(CMAX + 1);
	for (int32_t k = 0; k < n; ++k) {
		if (orf_off[k] < 0 || orf_len[k] < 0 || orf_off[k] + orf_len[k] > l_aa) { nsh_set_error("ps_gapped: item %d out of range", k); return 0; }
		if (full && orf_len[k] > 64 * CMAX) { nsh_set_error("ps_gapped: item %d: full-matrix ORF of %d aa exceeds %d", k, orf_len[k], 64 * CMAX); return 0; }
		for (int32_t t = 0; t < orf_len[k]; ++t)
			if (aa[orf_off[k] + t] >= 22) { nsh_set_error("ps_gapped: item %d: residue code %d", k, aa[orf_off[k] + t]); return 0; }
		const int w = full? orf_len[k] : 2 * band + 1;
		by_c[std::max(1, (w + 63) / 64)].push_back(k);
	}
	std::lock_guard<std::mutex> lk(g_mu);
	static hipStream_t stream = nullptr;
	int8_t *d_pssm = nullptr; uint8_t *d_aa = nullptr; int64_t *d_off = nullptr;
	int32_t *d_len = nullptr, *d_diag = nullptr, *d_items = nullptr, *d_out = nullptr;
	int ok = 0;
	const auto t0 = std::chrono::steady_clock::now();
	{
		if (!stream) PGOK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		PGOK(hipMalloc((void**)&d_pssm, (size_t)L * 22 + 4));
		PGOK(hipMalloc((void**)&d_aa, (size_t)l_aa + 4));
		PGOK(hipMalloc((void**)&d_off, (size_t)n * 8));
		PGOK(hipMalloc((void**)&d_len, (size_t)n * 4));
		PGOK(hipMalloc((void**)&d_diag, (size_t)n * 4));
		PGOK(hipMalloc((void**)&d_items, (size_t)n * 4));
		PGOK(hipMalloc((void**)&d_out, (size_t)n * 4));
		if (L) PGOK(hipMemcpyAsync(d_pssm, pssm, (size_t)L * 22, hipMemcpyHostToDevice, stream));
		if (l_aa) PGOK(hipMemcpyAsync(d_aa, aa, (size_t)l_aa, hipMemcpyHostToDevice, stream));
		PGOK(hipMemcpyAsync(d_off, orf_off, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PGOK(hipMemcpyAsync(d_len, orf_len, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		PGOK(hipMemcpyAsync(d_diag, diag, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		std::vector<int32_t> items;
		for (int c = 1; c <= CMAX; ++c) items.insert(items.end(), by_c[c].begin(), by_c[c].end());
		PGOK(hipMemcpyAsync(d_items, items.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream));
		int32_t first = 0;
		for (int c = 1; c <= CMAX; ++c) {
			if (by_c[c].empty()) continue;
			GapArgs ga;
			ga.pssm = d_pssm, ga.L = L, ga.prow = nullptr, ga.plen = nullptr, ga.aa = d_aa, ga.off = d_off, ga.len = d_len, ga.diag = d_diag;
			ga.band = band, ga.go = go, ga.ge = ge;
			ga.items = d_items + first, ga.n_items = (int32_t)by_c[c].size(), ga.out = d_out;
			launch(c, ga, full, stream);
			PGOK(hipGetLastError());
			first += ga.n_items;
			nsh_count_launch(1);
		}
		PGOK(hipMemcpyAsync(sw, d_out, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
		PGOK(hipStreamSynchronize(stream));
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_pssm, d_aa, d_off, d_len, d_diag, d_items, d_out };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}
