// A read of known size through the scalar data cache, for scaling hardware counter readings
// (rocprofv3 FETCH_SIZE) by access path: the extension DP reads its row records with scalar loads
// (64-byte lines), while a runtime device-to-device copy, the usual calibration, reads through the
// vector memory path.
#include <hip/hip_runtime.h>
#include "nsh_internal.h"

namespace {
typedef const uint32_t __attribute__((address_space(4))) *cu32_t;
constexpr int64_t kWordsPerWave = 16384;   // 64 KiB per wave

// wave w reads words [w * 16384, (w + 1) * 16384) with wave-uniform addresses (scalar loads),
// folds them into one word and writes it with a vector store from lane 0
__global__ void __launch_bounds__(64) scalar_read_kernel(const uint32_t *src, uint32_t *out)
{
	const int64_t w0 = (int64_t)blockIdx.x * kWordsPerWave;
	uint32_t x = 0;
	for (int64_t i = w0; i < w0 + kWordsPerWave; i += 16) {
#pragma unroll
		for (int k = 0; k < 16; ++k) x ^= ((cu32_t)src)[i + k];
	}
	if (threadIdx.x == 0) out[blockIdx.x] = x;
}
}   // namespace

extern "C" int nsh_diag_scalar_read(const void *src, int64_t bytes, uint32_t *out)
{
	if (!src || !out || bytes <= 0 || bytes % (kWordsPerWave * 4)) {
		nsh_set_error("nsh_diag_scalar_read: bytes must be a positive multiple of %lld", (long long)(kWordsPerWave * 4));
		return 0;
	}
	if (!nsh_ensure_device()) return 0;
	const int64_t waves = bytes / (kWordsPerWave * 4);
	if (waves > INT32_MAX) { nsh_set_error("nsh_diag_scalar_read: too many bytes"); return 0; }
	hipLaunchKernelGGL(scalar_read_kernel, dim3((uint32_t)waves), dim3(64), 0, 0, (const uint32_t*)src, out);
	const hipError_t e = hipDeviceSynchronize();
	if (e != hipSuccess || hipGetLastError() != hipSuccess) {
		nsh_set_error("nsh_diag_scalar_read: %s", hipGetErrorString(e));
		return 0;
	}
	return 1;
}
