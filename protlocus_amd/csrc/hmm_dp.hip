// Profile-HMM Viterbi scoring of DNA regions (mp_hmm_dp_score semantics, hmm.c:252-291; GPU
// entry points of hmm-gpu.c:282-352).
//
// The recurrence has no dependency inside a row: every state j of codon row i reads only row
// i-1 (states j-1 and j). One wave64 scores one region: lane L owns states [L*C+1, L*C+C]
// (C = ceil(M/64) <= 16), keeps pm/pi/pd and the seven transition scores of its states in
// registers, and takes the previous row's values of state L*C from lane L-1 by a DPP shift
// (lane 0 reads state 0, which is always 0). The codon of each row is translated once per 64
// rows (lane l translates row base+l, then the row's residue is read with readlane); the
// emission of the row comes from an LDS table emitT[residue][state] shared by the work-group,
// with residue 20 (stop codon or a base outside ACGT) an all-zero row. Each lane keeps a running
// maximum of cm and ci; one reduction per region gives the score. Work-groups are persistent
// and take regions, longest first, from a queue.
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>
#include "nsh_internal.h"

namespace {

#define DEV static __device__ __forceinline__

constexpr int kWaves = 4;       // regions in flight per work-group
constexpr int kCmax = 16;       // states per lane: M <= 1024

struct HmmArgs {
	const uint8_t *nt;          // concatenated regions, base codes (0..3, anything else = no emission)
	const int64_t *off;         // region offsets
	const int32_t *len;         // region lengths (nt)
	const int32_t *order;       // regions, longest first
	int32_t n;
	const int16_t *emitT;       // [21][64*C] emission by residue (HMMER column order) and state
	const int16_t *trans;       // [64*C][7] (zero past M)
	const uint8_t *calpha;      // codon -> HMMER column, 20 = none
	int32_t M, C;
	int32_t *counter;
	int32_t *score;
};

__global__ void fill_i32(int32_t *p, int32_t v, int32_t n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

template <int CTRL> DEV int dpp_in0(int x) { return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false); }
constexpr int WAVE_SHR1 = 0x138;

template <int C>
DEV int32_t hmm_region(const HmmArgs &a, const int16_t *emitT, const uint8_t *nt, int32_t nlen, int L)
{
	const int W = 64 * C;
	int t[C][7];
#pragma unroll
	for (int k = 0; k < C; ++k)
#pragma unroll
		for (int r = 0; r < 7; ++r) t[k][r] = a.trans[(L * C + k) * 7 + r];
	// states past M have all transitions 0 and emissions 0; they only feed states to their right,
	// which do not exist, and their own cm/ci are masked out of the running maximum
	bool live[C];
#pragma unroll
	for (int k = 0; k < C; ++k) live[k] = L * C + k < a.M;
	int pm[C], pi[C], pd[C];
#pragma unroll
	for (int k = 0; k < C; ++k) pm[k] = pi[k] = pd[k] = 0;
	int best = 0;
	const int rows = nlen / 3;
	int res = 20;
	for (int i = 0; i < rows; ++i) {
		if ((i & 63) == 0) {   // translate the codons of rows i..i+63, one per lane
			const int r = i + L;
			res = 20;
			if (r < rows) {
				const int b0 = nt[3 * r], b1 = nt[3 * r + 1], b2 = nt[3 * r + 2];
				if (b0 < 4 && b1 < 4 && b2 < 4) res = a.calpha[b0 << 4 | b1 << 2 | b2];
			}
		}
		const int e_row = __builtin_amdgcn_readlane(res, i & 63);
		const int16_t *er = emitT + e_row * W + L * C;
		int em[C];
#pragma unroll
		for (int k = 0; k < C; ++k) em[k] = er[k];
		// row i-1 values of state L*C (the lane's left neighbour); state 0 is 0
		const int lm = dpp_in0<WAVE_SHR1>(pm[C - 1]), li = dpp_in0<WAVE_SHR1>(pi[C - 1]), ld = dpp_in0<WAVE_SHR1>(pd[C - 1]);
		int cm[C], ci[C], cd[C];
#pragma unroll
		for (int k = 0; k < C; ++k) {
			const int m1 = k? pm[k - 1] : lm, i1 = k? pi[k - 1] : li, d1 = k? pd[k - 1] : ld;
			const int from = max(max(m1 + t[k][0], i1 + t[k][3]), d1 + t[k][5]);
			cm[k] = max(from + em[k], 0);
			ci[k] = max(max(pm[k] + t[k][1], pi[k] + t[k][4]), 0);
			cd[k] = max(max(m1 + t[k][2], d1 + t[k][6]), 0);
			if (live[k]) best = max(best, max(cm[k], ci[k]));
		}
#pragma unroll
		for (int k = 0; k < C; ++k) pm[k] = cm[k], pi[k] = ci[k], pd[k] = cd[k];
	}
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) best = max(best, __shfl_xor(best, d));
	return best;
}

template <int C>
__global__ void __launch_bounds__(64 * kWaves) hmm_kernel(HmmArgs a)
{
	extern __shared__ int16_t emitT[];
	const int L = threadIdx.x & 63;
	const int W = 64 * C;
	for (int k = threadIdx.x; k < 21 * W; k += blockDim.x) emitT[k] = a.emitT[k];
	__syncthreads();
	// The queue loop has no lane-dependent branch: every lane takes part in the atomic (lane 0 adds
	// 1, the others 0) and the ticket is lane 0's value; every lane stores the (wave-uniform)
	// score. A `lane == 0` branch at the bottom of the body next to one at the top may be merged
	// across the back-edge into a loop that lanes 1-63 re-enter without lane 0.
	for (;;) {
		const int q = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, L == 0? 1 : 0));
		if (q >= a.n) break;
		const int r = a.order[q];
		a.score[r] = hmm_region<C>(a, emitT, a.nt + a.off[r], a.len[r], L);
	}
}

using KernelFn = void (*)(HmmArgs);
template <int... Cs> struct Table { static constexpr KernelFn fn[sizeof...(Cs)] = { &hmm_kernel<Cs>... }; };
using KTab = Table<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

} // namespace

// A profile resident on the device: transposed emissions, transitions, codon map.
struct nsh_hmm_s {
	int32_t M = 0, C = 0;
	int16_t *d_emitT = nullptr, *d_trans = nullptr;
	uint8_t *d_calpha = nullptr;
	int32_t *d_counter = nullptr;
	hipStream_t stream = nullptr;
	std::mutex mu;
	// per-batch device buffers, grown on demand
	uint8_t *d_nt = nullptr; size_t cap_nt = 0;
	int64_t *d_off = nullptr; int32_t *d_len = nullptr, *d_order = nullptr, *d_score = nullptr; size_t cap_n = 0;
	int ncu = 0;
	hipEvent_t e0 = nullptr, e1 = nullptr;   // around the last Viterbi launch
	float last_ms = 0.f;
};

#define HMMOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("hmm: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

// HMMER column order of the 20 residues, and the codon -> column map (hmm.c:235-250)
static void hmm_codon_alpha(uint8_t tab[64])
{
	static const char alpha[] = "ACDEFGHIKLMNPQRSTVWY";
	int8_t a2a[22];
	for (int i = 0; i < 22; ++i) a2a[i] = -1;
	for (int i = 0; i < 20; ++i) a2a[ns_tab_aa20[(uint8_t)alpha[i]]] = (int8_t)i;
	for (int c = 0; c < 64; ++c) {
		const int r = ns_tab_codon[c];
		tab[c] = (uint8_t)(r < 20 && a2a[r] >= 0? a2a[r] : 20);
	}
}

extern "C" nsh_hmm_t *nsh_hmm_create(int32_t M, const int16_t *emit, const int16_t *trans)
{
	if (!nsh_ensure_device()) return nullptr;
	if (M < 1 || M > 64 * kCmax || !emit || !trans) { nsh_set_error("hmm: profile length %d outside 1..%d", M, 64 * kCmax); return nullptr; }
	nsh_hmm_t *h = new nsh_hmm_s();
	h->M = M, h->C = (M + 63) / 64;
	const int W = 64 * h->C;
	std::vector<int16_t> et((size_t)21 * W, 0), tr((size_t)W * 7, 0);
	for (int j = 0; j < M; ++j) {
		for (int r = 0; r < 20; ++r) et[(size_t)r * W + j] = emit[(size_t)j * 20 + r];
		for (int r = 0; r < 7; ++r) tr[(size_t)j * 7 + r] = trans[(size_t)j * 7 + r];
	}
	uint8_t calpha[64];
	hmm_codon_alpha(calpha);
	int dev = 0;
	HMMOK(hipGetDevice(&dev));
	HMMOK(hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev));
	HMMOK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
	HMMOK(hipEventCreate(&h->e0));
	HMMOK(hipEventCreate(&h->e1));
	HMMOK(hipMalloc((void**)&h->d_emitT, et.size() * 2));
	HMMOK(hipMalloc((void**)&h->d_trans, tr.size() * 2));
	HMMOK(hipMalloc((void**)&h->d_calpha, 64));
	HMMOK(hipMalloc((void**)&h->d_counter, 4));
	HMMOK(hipMemcpy(h->d_emitT, et.data(), et.size() * 2, hipMemcpyHostToDevice));
	HMMOK(hipMemcpy(h->d_trans, tr.data(), tr.size() * 2, hipMemcpyHostToDevice));
	HMMOK(hipMemcpy(h->d_calpha, calpha, 64, hipMemcpyHostToDevice));
	return h;
fail:
	nsh_hmm_destroy(h);
	return nullptr;
}

extern "C" float nsh_hmm_last_ms(nsh_hmm_t *h) { return h? h->last_ms : 0.f; }

extern "C" void nsh_hmm_destroy(nsh_hmm_t *h)
{
	if (!h) return;
	void *bufs[] = { h->d_emitT, h->d_trans, h->d_calpha, h->d_counter, h->d_nt, h->d_off, h->d_len, h->d_order, h->d_score };
	for (void *b : bufs) if (b) (void)hipFree(b);
	if (h->e0) (void)hipEventDestroy(h->e0);
	if (h->e1) (void)hipEventDestroy(h->e1);
	if (h->stream) (void)hipStreamDestroy(h->stream);
	delete h;
}

extern "C" int nsh_hmm_score_batch(nsh_hmm_t *h, int32_t n, const uint8_t *nt, int64_t l_nt, const int32_t *off,
		const int32_t *len, int32_t *score)
{
	if (!h || n < 0 || (n > 0 && (!nt || !off || !len || !score))) { nsh_set_error("hmm: bad batch arguments"); return 0; }
	if (n == 0) return 1;
	for (int32_t i = 0; i < n; ++i)
		if (off[i] < 0 || len[i] < 0 || (int64_t)off[i] + len[i] > l_nt) { nsh_set_error("hmm: region %d out of range", i); return 0; }
	std::lock_guard<std::mutex> lk(h->mu);
	std::vector<int32_t> order(n);
	std::vector<int64_t> off64(off, off + n);
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return len[x] > len[y]; });
	const auto t0 = std::chrono::steady_clock::now();
	int ok = 0;
	{
		if ((size_t)l_nt + 1 > h->cap_nt) {
			if (h->d_nt) (void)hipFree(h->d_nt), h->d_nt = nullptr, h->cap_nt = 0;
			HMMOK(hipMalloc((void**)&h->d_nt, (size_t)l_nt + 1));
			h->cap_nt = (size_t)l_nt + 1;
		}
		if ((size_t)n > h->cap_n) {
			void *bufs[] = { h->d_off, h->d_len, h->d_order, h->d_score };
			for (void *b : bufs) if (b) (void)hipFree(b);
			h->d_off = nullptr, h->d_len = h->d_order = h->d_score = nullptr, h->cap_n = 0;
			HMMOK(hipMalloc((void**)&h->d_off, (size_t)n * 8));
			HMMOK(hipMalloc((void**)&h->d_len, (size_t)n * 4));
			HMMOK(hipMalloc((void**)&h->d_order, (size_t)n * 4));
			HMMOK(hipMalloc((void**)&h->d_score, (size_t)n * 4));
			h->cap_n = n;
		}
		HMMOK(hipMemcpyAsync(h->d_nt, nt, (size_t)l_nt, hipMemcpyHostToDevice, h->stream));
		HMMOK(hipMemcpyAsync(h->d_off, off64.data(), (size_t)n * 8, hipMemcpyHostToDevice, h->stream));
		HMMOK(hipMemcpyAsync(h->d_len, len, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
		HMMOK(hipMemcpyAsync(h->d_order, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
		hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, h->stream, h->d_counter, 0, 1);
		HmmArgs a;
		a.nt = h->d_nt, a.off = h->d_off, a.len = h->d_len, a.order = h->d_order, a.n = n;
		a.emitT = h->d_emitT, a.trans = h->d_trans, a.calpha = h->d_calpha, a.M = h->M, a.C = h->C;
		a.counter = h->d_counter, a.score = h->d_score;
		const int grid = std::max(1, std::min((n + kWaves - 1) / kWaves, h->ncu * 4));
		const size_t lds = (size_t)21 * 64 * h->C * 2;
		HMMOK(hipEventRecord(h->e0, h->stream));
		hipLaunchKernelGGL(KTab::fn[h->C - 1], dim3(grid), dim3(64 * kWaves), lds, h->stream, a);
		HMMOK(hipGetLastError());
		HMMOK(hipEventRecord(h->e1, h->stream));
		HMMOK(hipMemcpyAsync(score, h->d_score, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream));
		HMMOK(hipStreamSynchronize(h->stream));
		HMMOK(hipEventElapsedTime(&h->last_ms, h->e0, h->e1));
		nsh_count_launch(2);
		ok = 1;
	}
fail:
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}

// ---------------------------------------------------------------------------------------
// the reference's GPU entry points (mppriv.h:128-134). The profile object and its accessors
// belong to the caller's hmm.c. They are weak references here, bound when the caller's program is
// linked against this library (the linker then exports the caller's definitions); a caller that
// loads its accessors after this library (a plugin, a test harness) is found by name in the
// process's global scope instead. Without them the entry points fail and the caller keeps its CPU
// path. The device copy of a profile is cached in the caller's slot 0 and freed by
// mp_hmm_gpu_release() (called from mp_hmm_destroy).
// ---------------------------------------------------------------------------------------
typedef struct mp_hmm_s mp_hmm_t;   // the caller's profile type (hmm.c)
extern "C" {
__attribute__((weak)) int32_t mp_hmm_length(const mp_hmm_t *h);
__attribute__((weak)) const int16_t *mp_hmm_emissions(const mp_hmm_t *h);
__attribute__((weak)) const int16_t *mp_hmm_transitions(const mp_hmm_t *h);
__attribute__((weak)) void **mp_hmm_gpu_slot(const mp_hmm_t *h, int which);
}

static std::mutex g_slot_mu;

namespace {
struct HmmAccess {
	int32_t (*length)(const mp_hmm_t*);
	const int16_t *(*emissions)(const mp_hmm_t*);
	const int16_t *(*transitions)(const mp_hmm_t*);
	void **(*slot)(const mp_hmm_t*, int);
};
// the caller's accessors: the link-time references, else a lookup by name (once found, kept)
bool hmm_access(HmmAccess &a)
{
	static HmmAccess found = {};
	static std::mutex mu;
	std::lock_guard<std::mutex> lk(mu);
	if (!found.length) {
		HmmAccess t = { mp_hmm_length, mp_hmm_emissions, mp_hmm_transitions, mp_hmm_gpu_slot };
		if (!t.length || !t.emissions || !t.transitions || !t.slot) {
			t.length = (int32_t (*)(const mp_hmm_t*))dlsym(RTLD_DEFAULT, "mp_hmm_length");
			t.emissions = (const int16_t *(*)(const mp_hmm_t*))dlsym(RTLD_DEFAULT, "mp_hmm_emissions");
			t.transitions = (const int16_t *(*)(const mp_hmm_t*))dlsym(RTLD_DEFAULT, "mp_hmm_transitions");
			t.slot = (void **(*)(const mp_hmm_t*, int))dlsym(RTLD_DEFAULT, "mp_hmm_gpu_slot");
		}
		if (t.length && t.emissions && t.transitions && t.slot) found = t;
	}
	a = found;
	return found.length != nullptr;
}
} // namespace

static nsh_hmm_t *hmm_of(const mp_hmm_t *h)
{
	HmmAccess acc;
	if (!hmm_access(acc)) {
		nsh_set_error("hmm: the caller's profile accessors (hmm.c) are not linked");
		return nullptr;
	}
	void **slot = acc.slot(h, 0);
	if (!slot) return nullptr;
	std::lock_guard<std::mutex> lk(g_slot_mu);
	if (!*slot) *slot = nsh_hmm_create(acc.length(h), acc.emissions(h), acc.transitions(h));
	return (nsh_hmm_t*)*slot;
}

extern "C" int mp_hmm_dp_score_gpu_batch(const mp_hmm_t *h, int32_t n, const uint8_t *nt, int64_t l_nt,
		const int32_t *off, const int32_t *len, int32_t *score)
{
	if (!h || n <= 0 || !nt || l_nt <= 0) return 0;
	nsh_hmm_t *g = hmm_of(h);
	return g? nsh_hmm_score_batch(g, n, nt, l_nt, off, len, score) : 0;
}

extern "C" int32_t mp_hmm_dp_score_gpu(const mp_hmm_t *h, const uint8_t *nt, int32_t nlen)
{
	int32_t off = 0, score;
	if (nlen < 3) return INT32_MIN;
	return mp_hmm_dp_score_gpu_batch(h, 1, nt, nlen, &off, &nlen, &score)? score : INT32_MIN;
}

extern "C" void mp_hmm_gpu_release(mp_hmm_t *h)
{
	HmmAccess acc;
	if (!h || !hmm_access(acc)) return;
	void **slot = acc.slot(h, 0);
	if (!slot) return;
	std::lock_guard<std::mutex> lk(g_slot_mu);
	if (*slot) nsh_hmm_destroy((nsh_hmm_t*)*slot), *slot = nullptr;
}
