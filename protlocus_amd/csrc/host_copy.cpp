// Large copies between pageable host memory and the device, for the entry points that take the
// caller's host buffers (anchors, chains, genomes, index tables).
//
// A pageable hipMemcpy of a gigabyte moves 7-10 GB/s here. These copies go through two pinned
// 64-MB staging buffers instead: while one chunk crosses PCIe by DMA, the host copies the chunk
// before it between the caller's buffer and the other staging buffer with several threads. The
// staging buffers are allocated on the first large copy and kept for the process's lifetime
// (128 MB of pinned memory); a mutex serialises their users. When they cannot be allocated (or
// NSH_NO_PINNED_STAGING=1 is set), a copy is one hipMemcpyAsync on the caller's stream followed by
// a wait on that stream, so it stays ordered after the kernels queued there.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
#include "nsh_internal.h"

namespace {

constexpr size_t kChunk = (size_t)64 << 20;   // staging chunk
constexpr size_t kDirect = (size_t)32 << 20;  // below this a plain copy is as fast
constexpr int kThreads = 8;                    // host threads per chunk copy

std::mutex g_mu;
void *g_pin[2] = {};

bool ensure_pinned()
{
	const char *off = std::getenv("NSH_NO_PINNED_STAGING");   // read per copy, so a test can switch it
	if (off && std::atoi(off)) return false;
	for (int b = 0; b < 2; ++b)
		if (!g_pin[b] && hipHostMalloc(&g_pin[b], kChunk, hipHostMallocDefault) != hipSuccess) {
			g_pin[b] = nullptr;
			return false;
		}
	return true;
}

void par_memcpy(void *dst, const void *src, size_t n)
{
	if (n < ((size_t)4 << 20)) { std::memcpy(dst, src, n); return; }
	const size_t part = (n / kThreads + 4095) & ~(size_t)4095;
	std::vector<std::thread> th;
	for (int t = 1; t < kThreads; ++t) {
		const size_t o = (size_t)t * part;
		if (o >= n) break;
		th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, std::min(part, n - o)); });
	}
	std::memcpy(dst, src, std::min(part, n));
	for (auto &x : th) x.join();
}

// one copy on stream s, then a wait on s (the copy follows the kernels already queued on s)
int plain_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s, const char *what)
{
	hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (e != hipSuccess) { nsh_set_error("%s of %zu bytes failed: %s", what, bytes, hipGetErrorString(e)); return 0; }
	return 1;
}

} // namespace

int nsh_copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s)
{
	if (bytes == 0) return 1;
	if (bytes < kDirect) return plain_copy(dst, src, bytes, hipMemcpyDeviceToHost, s, "copy to host");
	std::lock_guard<std::mutex> lk(g_mu);
	if (!ensure_pinned()) return plain_copy(dst, src, bytes, hipMemcpyDeviceToHost, s, "copy to host");
	hipEvent_t ev[2] = {};
	int ok = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
			hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
	const size_t nch = (bytes + kChunk - 1) / kChunk;
	auto issue = [&](size_t k) {
		const size_t o = k * kChunk, n = std::min(kChunk, bytes - o);
		return hipMemcpyAsync(g_pin[k & 1], (const char*)src + o, n, hipMemcpyDeviceToHost, s) == hipSuccess &&
				hipEventRecord(ev[k & 1], s) == hipSuccess;
	};
	if (ok) ok = issue(0);
	for (size_t k = 0; ok && k < nch; ++k) {
		// the buffer of chunk k+1 held chunk k-1, already copied out
		if (k + 1 < nch) ok = issue(k + 1);
		if (ok) ok = hipEventSynchronize(ev[k & 1]) == hipSuccess;
		if (ok) {
			const size_t o = k * kChunk;
			par_memcpy((char*)dst + o, g_pin[k & 1], std::min(kChunk, bytes - o));
		}
	}
	(void)hipStreamSynchronize(s);
	for (auto e : ev) if (e) (void)hipEventDestroy(e);
	if (!ok) nsh_set_error("copy to host failed");
	return ok;
}

int nsh_copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t s)
{
	if (bytes == 0) return 1;
	if (bytes < kDirect) {
		const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
		if (e != hipSuccess) { nsh_set_error("copy to device of %zu bytes failed: %s", bytes, hipGetErrorString(e)); return 0; }
		return 1;
	}
	std::lock_guard<std::mutex> lk(g_mu);
	if (!ensure_pinned()) return plain_copy(dst, src, bytes, hipMemcpyHostToDevice, s, "copy to device");
	hipEvent_t ev[2] = {};
	int ok = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
			hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
	const size_t nch = (bytes + kChunk - 1) / kChunk;
	for (size_t k = 0; ok && k < nch; ++k) {
		const size_t o = k * kChunk, n = std::min(kChunk, bytes - o);
		if (k >= 2) ok = hipEventSynchronize(ev[k & 1]) == hipSuccess;   // chunk k-2 has left this buffer
		if (!ok) break;
		par_memcpy(g_pin[k & 1], (const char*)src + o, n);
		ok = hipMemcpyAsync((char*)dst + o, g_pin[k & 1], n, hipMemcpyHostToDevice, s) == hipSuccess &&
				hipEventRecord(ev[k & 1], s) == hipSuccess;
	}
	// the staging buffers are reused by the next copy: wait until both have been read
	(void)hipStreamSynchronize(s);
	for (auto e : ev) if (e) (void)hipEventDestroy(e);
	if (!ok) nsh_set_error("copy to device failed");
	return ok;
}

#include <chrono>
#include <cstdio>
#include <cstdlib>

static double nsh_now_ms()
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

nsh_step_clock::nsh_step_clock(const char *w) : who(w), on(std::getenv("NSH_PROFILE") != nullptr), t(nsh_now_ms()) {}

void nsh_step_clock::lap(const char *what, int64_t n)
{
	if (!on) return;
	const double now = nsh_now_ms();
	if (n >= 0) std::fprintf(stderr, "[%s] %-24s %9.1f ms  n=%lld\n", who, what, now - t, (long long)n);
	else std::fprintf(stderr, "[%s] %-24s %9.1f ms\n", who, what, now - t);
	t = now;
}
