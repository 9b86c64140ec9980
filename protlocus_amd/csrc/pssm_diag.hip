// Best ungapped PSSM segment per (ORF, diagonal) (ps_ungapped, prescreen.c:334-344).
//
// One lane per (ORF, diagonal) item: a running sum over the positions of the diagonal, clamped at
// zero, with the first strict maximum recorded (query interval [qs, qe)). Items are taken longest
// first so the lanes of a wave run loops of similar length. The PSSM rows stay in HBM and are read
// through the caches (one profile per call, at most a few tens of KB).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <mutex>
#include <numeric>
#include <vector>
#include "nsh_internal.h"

namespace {

struct DiagArgs {
	const int8_t *pssm; int32_t L;
	const uint8_t *aa;
	const int64_t *off; const int32_t *len, *diag, *order;
	int32_t n;
	int32_t *out;   // 3 ints per item: score, qs, qe
};

__global__ void __launch_bounds__(256) pssm_diag_kernel(DiagArgs a)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= a.n) return;
	const int k = a.order[t];
	const uint8_t *orf = a.aa + a.off[k];
	const int n = a.len[k], diag = a.diag[k], L = a.L;
	const int i0 = diag < 0? -diag : 0, i1 = n - diag < L? n - diag : L;
	int sc = 0, best = 0, st = i0, qs = i0, qe = i0;
	for (int i = i0; i < i1; ++i) {
		const int r = orf[i + diag];
		sc += a.pssm[(int64_t)i * 22 + (r < 22? r : 21)];
		if (sc < 0) sc = 0, st = i + 1;
		else if (sc > best) best = sc, qs = st, qe = i + 1;
	}
	int32_t *o = a.out + 3 * (int64_t)k;
	o[0] = best, o[1] = qs, o[2] = qe;
}

std::mutex g_mu;

} // namespace

#define PDOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("pssm_diag: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

extern "C" int nsh_pssm_ungapped_batch(const int8_t *pssm, int32_t L, const uint8_t *aa, int64_t l_aa, int32_t n,
		const int64_t *orf_off, const int32_t *orf_len, const int32_t *diag, int32_t *score, int32_t *qs, int32_t *qe)
{
	if (!nsh_ensure_device()) return 0;
	if (!pssm || L < 1 || !aa || l_aa < 0 || n < 0 || (n > 0 && (!orf_off || !orf_len || !diag || !score || !qs || !qe))) {
		nsh_set_error("pssm_diag: bad arguments");
		return 0;
	}
	if (n == 0) return 1;
	std::vector<int32_t> order(n), out((size_t)n * 3);
	std::vector<int32_t> work(n);
	for (int32_t k = 0; k < n; ++k) {
		if (orf_off[k] < 0 || orf_len[k] < 0 || orf_off[k] + orf_len[k] > l_aa) { nsh_set_error("pssm_diag: item %d out of range", k); return 0; }
		const int32_t i0 = diag[k] < 0? -diag[k] : 0, i1 = orf_len[k] - diag[k] < L? orf_len[k] - diag[k] : L;
		work[k] = i1 > i0? i1 - i0 : 0;
	}
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return work[x] > work[y]; });
	std::lock_guard<std::mutex> lk(g_mu);
	static hipStream_t stream = nullptr;
	int8_t *d_pssm = nullptr; uint8_t *d_aa = nullptr; int64_t *d_off = nullptr;
	int32_t *d_len = nullptr, *d_diag = nullptr, *d_order = nullptr, *d_out = nullptr;
	int ok = 0;
	const auto t0 = std::chrono::steady_clock::now();
	{
		if (!stream) PDOK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		PDOK(hipMalloc((void**)&d_pssm, (size_t)L * 22));
		PDOK(hipMalloc((void**)&d_aa, (size_t)l_aa + 1));
		PDOK(hipMalloc((void**)&d_off, (size_t)n * 8));
		PDOK(hipMalloc((void**)&d_len, (size_t)n * 4));
		PDOK(hipMalloc((void**)&d_diag, (size_t)n * 4));
		PDOK(hipMalloc((void**)&d_order, (size_t)n * 4));
		PDOK(hipMalloc((void**)&d_out, (size_t)n * 12));
		PDOK(hipMemcpyAsync(d_pssm, pssm, (size_t)L * 22, hipMemcpyHostToDevice, stream));
		if (l_aa) PDOK(hipMemcpyAsync(d_aa, aa, (size_t)l_aa, hipMemcpyHostToDevice, stream));
		PDOK(hipMemcpyAsync(d_off, orf_off, (size_t)n * 8, hipMemcpyHostToDevice, stream));
		PDOK(hipMemcpyAsync(d_len, orf_len, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		PDOK(hipMemcpyAsync(d_diag, diag, (size_t)n * 4, hipMemcpyHostToDevice, stream));
		PDOK(hipMemcpyAsync(d_order, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream));
		DiagArgs a;
		a.pssm = d_pssm, a.L = L, a.aa = d_aa, a.off = d_off, a.len = d_len, a.diag = d_diag, a.order = d_order;
		a.n = n, a.out = d_out;
		hipLaunchKernelGGL(pssm_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a);
		PDOK(hipGetLastError());
		PDOK(hipMemcpyAsync(out.data(), d_out, (size_t)n * 12, hipMemcpyDeviceToHost, stream));
		PDOK(hipStreamSynchronize(stream));
		for (int32_t k = 0; k < n; ++k) score[k] = out[3 * k], qs[k] = out[3 * k + 1], qe[k] = out[3 * k + 2];
		nsh_count_launch(1);
		ok = 1;
	}
fail:
	{
		void *bufs[] = { d_pssm, d_aa, d_off, d_len, d_diag, d_order, d_out };
		for (void *b : bufs) if (b) (void)hipFree(b);
	}
	nsh_add_times(0, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), 0);
	return ok;
}
