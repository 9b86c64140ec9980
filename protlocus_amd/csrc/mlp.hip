// Learned chain-candidate prescreen: a 64 -> 256 -> 256 -> 1 MLP (ReLU, logistic output) with
// hand-written BF16 MFMA kernels for inference and training (FP32 accumulation, FP32 Adam
// master weights). No reference counterpart: it ranks candidates ahead of the spliced DP,
// which stays the fallback and the arbiter of recall.
//
// Matrix products use v_mfma_f32_32x32x16_bf16 (gfx950): a wave computes a 32 x 32 FP32 tile;
// lane l supplies row (l % 32) of A and column (l % 32) of B, eight K values each (the K
// positions of a lane are the same for A and B, so the contraction does not depend on how the
// hardware orders them); the 16 accumulators of lane l hold rows 8*(r/4) + 4*(l/32) + r%4,
// column l % 32.
//
// Layout: weights are kept as FP32 masters W[in][out] plus two BF16 copies, W[in][out] and its
// transpose WT[out][in], so every product reads its B operand with unit K stride in one of the
// two copies. Activations are BF16 row-major [sample][unit].
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>
#include "nsh_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;
#define DEV static __device__ __forceinline__

constexpr int IN = 64, HID = 256;

// float -> bf16 bits, round to nearest even (finite inputs)
DEV u16 f2bf(float f)
{
	unsigned u = __builtin_bit_cast(unsigned, f);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (u16)(u >> 16);
}
DEV float bf2f(u16 b) { return __builtin_bit_cast(float, (unsigned)b << 16); }

union V8 { bf16x8 v; u16 s[8]; uint4 q; };

// eight K-consecutive values of row `row` of a matrix with element (row, k) at p[row*sr + k*sk]
DEV bf16x8 load8(const u16 *p, long sr, long sk, int row, int k)
{
	V8 x;
	const u16 *q = p + (long)row * sr + (long)k * sk;
	if (sk == 1 && ((uintptr_t)q & 15) == 0) x.q = *(const uint4*)q;
	else {
#pragma unroll
		for (int e = 0; e < 8; ++e) x.s[e] = q[(long)e * sk];
	}
	return x.v;
}

// C[M,N] = A[M,K] x B[K,N], A(m,k) = A[m*sam + k*sak], B(k,n) = B[n*sbn + k*sbk]
struct GemmArgs {
	const u16 *A; long sam, sak;
	const u16 *B; long sbn, sbk;
	int M, N, K;
	int epi;                // 0: FP32 out; 1: bias + ReLU -> BF16; 2: x relu'(mask) -> BF16
	float *Cf; long scm;    // FP32 output (epi 0), row stride
	u16 *Cb; long sbm;      // BF16 output (epi 1, 2), row stride
	const float *bias;      // [N] (epi 1)
	const u16 *mask; long smm;   // activation whose sign gates the gradient (epi 2)
};

// one wave per 32 x 32 tile, four waves per work-group along N
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g)
{
	const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int m0 = blockIdx.x * 32, n0 = (blockIdx.y * 4 + w) * 32;
	if (n0 >= g.N) return;   // whole wave; no barrier in this kernel
	const int r = L & 31, kh = (L >> 5) * 8;
	f32x16 acc = {};
	for (int k0 = 0; k0 < g.K; k0 += 16) {
		const bf16x8 a = load8(g.A, g.sam, g.sak, m0 + r, k0 + kh);
		const bf16x8 b = load8(g.B, g.sbn, g.sbk, n0 + r, k0 + kh);
		acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
	}
	const int col = n0 + r;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int row = m0 + 8 * (i >> 2) + 4 * (L >> 5) + (i & 3);
		float v = acc[i];
		if (g.epi == 0) g.Cf[(long)row * g.scm + col] = v;
		else if (g.epi == 1) { v += g.bias[col]; g.Cb[(long)row * g.sbm + col] = f2bf(v > 0.f? v : 0.f); }
		else g.Cb[(long)row * g.sbm + col] = bf2f(g.mask[(long)row * g.smm + col]) > 0.f? f2bf(v) : (u16)0;
	}
}

// Fused inference: 64 samples per work-group (4 waves). Layer 1 and layer 2 run on MFMA with
// the activations in LDS; layer 3 (256 -> 1) is folded into layer 2's epilogue as a dot product
// with w3, reduced through LDS. Weights are read from HBM/L2 (WT copies, unit K stride).
__global__ void __launch_bounds__(256) mlp_infer_kernel(const u16 *X, int n, const u16 *W1T, const float *b1,
		const u16 *W2T, const float *b2, const float *w3, const float *b3, float *logit)
{
	__shared__ __attribute__((aligned(16))) u16 xs[64 * IN];     // 8 KB
	__shared__ __attribute__((aligned(16))) u16 h1[64 * HID];    // 32 KB
	__shared__ float part[4][64];
	const int L = threadIdx.x & 63, w = threadIdx.x >> 6, r = L & 31, kh = (L >> 5) * 8;
	const long s0 = (long)blockIdx.x * 64;
	// stage X (rows past n read as zero)
	for (int t = threadIdx.x; t < 64 * IN / 8; t += 256) {
		const int row = t / (IN / 8), c8 = t % (IN / 8);
		uint4 v = {0, 0, 0, 0};
		if (s0 + row < n) v = *(const uint4*)(X + (s0 + row) * IN + c8 * 8);
		*(uint4*)(xs + row * IN + c8 * 8) = v;
	}
	__syncthreads();
	// layer 1: wave w computes output columns [64w, 64w+64) for all 64 rows (2 x 2 tiles)
	for (int mt = 0; mt < 2; ++mt)
		for (int nt = 0; nt < 2; ++nt) {
			const int m0 = mt * 32, n0 = w * 64 + nt * 32;
			f32x16 acc = {};
#pragma unroll
			for (int k0 = 0; k0 < IN; k0 += 16) {
				const bf16x8 a = load8(xs, IN, 1, m0 + r, k0 + kh);
				const bf16x8 b = load8(W1T, IN, 1, n0 + r, k0 + kh);
				acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
			}
#pragma unroll
			for (int i = 0; i < 16; ++i) {
				const int row = m0 + 8 * (i >> 2) + 4 * (L >> 5) + (i & 3), col = n0 + r;
				const float v = acc[i] + b1[col];
				h1[row * HID + col] = f2bf(v > 0.f? v : 0.f);
			}
		}
	__syncthreads();
	// layer 2 + 3: per row, sum over this wave's 64 columns of relu(h2) * w3
	float dot[2][16];
	for (int mt = 0; mt < 2; ++mt)
#pragma unroll
		for (int i = 0; i < 16; ++i) dot[mt][i] = 0.f;
	for (int mt = 0; mt < 2; ++mt)
		for (int nt = 0; nt < 2; ++nt) {
			const int m0 = mt * 32, n0 = w * 64 + nt * 32;
			f32x16 acc = {};
			for (int k0 = 0; k0 < HID; k0 += 16) {
				const bf16x8 a = load8(h1, HID, 1, m0 + r, k0 + kh);
				const bf16x8 b = load8(W2T, HID, 1, n0 + r, k0 + kh);
				acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
			}
			const int col = n0 + r;
#pragma unroll
			for (int i = 0; i < 16; ++i) {
				float v = acc[i] + b2[col];
				v = bf2f(f2bf(v > 0.f? v : 0.f));   // the activation as stored in training
				dot[mt][i] += v * w3[col];
			}
		}
	// reduce each row's partial dot over the 32 columns held by lanes with the same L/32
	for (int mt = 0; mt < 2; ++mt)
#pragma unroll
		for (int i = 0; i < 16; ++i) {
			float v = dot[mt][i];
#pragma unroll
			for (int d = 16; d >= 1; d >>= 1) v += __shfl_xor(v, d);
			const int row = mt * 32 + 8 * (i >> 2) + 4 * (L >> 5) + (i & 3);
			if (r == 0) part[w][row] = v;
		}
	__syncthreads();
	if (threadIdx.x < 64 && s0 + threadIdx.x < n)
		logit[s0 + threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
			part[3][threadIdx.x] + b3[0];
}

// layer 3 forward for training: y = h2 . w3 + b3, dy = (sigmoid(y) - t) / B, loss terms
__global__ void head_kernel(const u16 *H2, const float *w3, const float *b3, const float *t, int B, float *y,
		float *dy, float *loss)
{
	const int s = blockIdx.x * 4 + (threadIdx.x >> 6), L = threadIdx.x & 63;
	if (s >= B) return;
	float v = 0.f;
	for (int c = L; c < HID; c += 64) v += bf2f(H2[(long)s * HID + c]) * w3[c];
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
	v += b3[0];
	if (L == 0) {
		y[s] = v;
		const float p = 1.f / (1.f + expf(-v));
		dy[s] = (p - t[s]) / (float)B;
		// BCE with logits, stable form
		loss[s] = fmaxf(v, 0.f) - v * t[s] + log1pf(expf(-fabsf(v)));
	}
}

constexpr int ROWS_PER_BLOCK = 16;

// dH2[s][c] = dy[s] w3[c] [H2 > 0] (BF16); dW3[c] += sum_s H2[s][c] dy[s] (zeroed beforehand).
// 16 samples per work-group, one column per thread.
__global__ void __launch_bounds__(256) head_back_kernel(const u16 *H2, const float *dy, const float *w3, int B,
		float *dW3, u16 *dH2)
{
	const int c = threadIdx.x, s0 = blockIdx.x * ROWS_PER_BLOCK;
	float acc = 0.f;
	for (int s = s0; s < s0 + ROWS_PER_BLOCK && s < B; ++s) {
		const float h = bf2f(H2[(long)s * HID + c]);
		acc += h * dy[s];
		dH2[(long)s * HID + c] = h > 0.f? f2bf(dy[s] * w3[c]) : (u16)0;
	}
	atomicAdd(&dW3[c], acc);
}

// out[c] += column sums of a BF16 [B][N] matrix over 16 rows per work-group (bias gradients)
__global__ void __launch_bounds__(256) colsum_kernel(const u16 *D, int B, int N, float *out)
{
	const int c = threadIdx.x, s0 = blockIdx.x * ROWS_PER_BLOCK;
	if (c >= N) return;
	float acc = 0.f;
	for (int s = s0; s < s0 + ROWS_PER_BLOCK && s < B; ++s) acc += bf2f(D[(long)s * N + c]);
	atomicAdd(&out[c], acc);
}
__global__ void fill_f32_kernel(float *p, long n, float v)
{
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
__global__ void vecsum_kernel(const float *v, int n, float *out)
{
	__shared__ float sm[256];
	float a = 0.f;
	for (int i = threadIdx.x; i < n; i += 256) a += v[i];
	sm[threadIdx.x] = a;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[0] = sm[0];
}

// Adam over a flat parameter vector; refreshes the BF16 copies of weight matrices
struct AdamArgs {
	float *p; const float *g; float *m, *v; int n;
	float lr, b1, b2, eps, c1, c2;   // c1 = 1 - b1^t, c2 = 1 - b2^t
};
__global__ void adam_kernel(AdamArgs a)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n) return;
	const float g = a.g[i];
	const float m = a.b1 * a.m[i] + (1.f - a.b1) * g, v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
	a.m[i] = m, a.v[i] = v;
	a.p[i] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
}
// W[in][out] (FP32) -> BF16 W[in][out] and WT[out][in]
__global__ void to_bf16_kernel(const float *W, int in, int out, u16 *Wb, u16 *WTb)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= in * out) return;
	const int r = i / out, c = i % out;
	const u16 b = f2bf(W[i]);
	Wb[i] = b;
	WTb[(long)c * in + r] = b;
}
// feature standardisation fused with the BF16 conversion: y = bf16((x - mu[f]) * isd[f]), f = i % 64
__global__ void f32_to_bf16_kernel(const float *x, long n, const float *mu, const float *isd, u16 *y)
{
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) { const int f = (int)(i % IN); y[i] = f2bf((x[i] - mu[f]) * isd[f]); }
}

} // namespace

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
// parameter vector layout (FP32): W1[64][256], b1[256], W2[256][256], b2[256], w3[256], b3[1]
enum : long { P_W1 = 0, P_B1 = P_W1 + IN * HID, P_W2 = P_B1 + HID, P_B2 = P_W2 + HID * HID, P_W3 = P_B2 + HID,
	P_B3 = P_W3 + HID, P_N = P_B3 + 1 };

struct nsh_mlp_s {
	hipStream_t stream = nullptr;
	float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // P_N each
	u16 *W1b = nullptr, *W1Tb = nullptr, *W2b = nullptr, *W2Tb = nullptr;
	int step = 0;
	// training scratch for batches up to cap
	int cap = 0;
	u16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dH1 = nullptr, *dH2 = nullptr;
	float *Xt = nullptr;   // FP32 staging of the batch
	float *t = nullptr, *y = nullptr, *dy = nullptr, *loss = nullptr, *lsum = nullptr;
	// inference scratch
	long icap = 0;
	u16 *Xi = nullptr; float *Xf = nullptr, *li = nullptr;
	float *norm = nullptr;   // mu[64], 1/sd[64] of the input standardisation (0 and 1 by default)
	std::mutex mu;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	float last_ms = 0.f;
};

#define MLPOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("mlp: %s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

static void gemm(hipStream_t s, const GemmArgs &g)
{
	hipLaunchKernelGGL(gemm_kernel, dim3(g.M / 32, (g.N / 32 + 3) / 4), dim3(256), 0, s, g);
}

static void refresh_bf16(nsh_mlp_t *h)
{
	hipLaunchKernelGGL(to_bf16_kernel, dim3((IN * HID + 255) / 256), dim3(256), 0, h->stream, h->p + P_W1, IN, HID, h->W1b, h->W1Tb);
	hipLaunchKernelGGL(to_bf16_kernel, dim3((HID * HID + 255) / 256), dim3(256), 0, h->stream, h->p + P_W2, HID, HID, h->W2b, h->W2Tb);
}

extern "C" nsh_mlp_t *nsh_mlp_create(const float *params)
{
	if (!nsh_ensure_device()) return nullptr;
	if (!params) { nsh_set_error("mlp: parameters required"); return nullptr; }
	nsh_mlp_t *h = new nsh_mlp_s();
	MLPOK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
	MLPOK(hipEventCreate(&h->e0));
	MLPOK(hipEventCreate(&h->e1));
	MLPOK(hipMalloc((void**)&h->p, P_N * 4));
	MLPOK(hipMalloc((void**)&h->g, P_N * 4));
	MLPOK(hipMalloc((void**)&h->m, P_N * 4));
	MLPOK(hipMalloc((void**)&h->v, P_N * 4));
	MLPOK(hipMalloc((void**)&h->W1b, IN * HID * 2));
	MLPOK(hipMalloc((void**)&h->W1Tb, IN * HID * 2));
	MLPOK(hipMalloc((void**)&h->W2b, HID * HID * 2));
	MLPOK(hipMalloc((void**)&h->W2Tb, HID * HID * 2));
	MLPOK(hipMalloc((void**)&h->norm, 2 * IN * 4));
	{
		float nm[2 * IN];
		for (int f = 0; f < IN; ++f) nm[f] = 0.f, nm[IN + f] = 1.f;
		MLPOK(hipMemcpy(h->norm, nm, sizeof(nm), hipMemcpyHostToDevice));
	}
	MLPOK(hipMemcpy(h->p, params, P_N * 4, hipMemcpyHostToDevice));
	MLPOK(hipMemset(h->m, 0, P_N * 4));
	MLPOK(hipMemset(h->v, 0, P_N * 4));
	refresh_bf16(h);
	MLPOK(hipGetLastError());
	MLPOK(hipStreamSynchronize(h->stream));
	return h;
fail:
	nsh_mlp_destroy(h);
	return nullptr;
}

extern "C" void nsh_mlp_destroy(nsh_mlp_t *h)
{
	if (!h) return;
	void *bufs[] = { h->p, h->g, h->m, h->v, h->W1b, h->W1Tb, h->W2b, h->W2Tb, h->X, h->H1, h->H2, h->dH1, h->dH2,
		h->norm, h->Xt, h->t, h->y, h->dy, h->loss, h->lsum, h->Xi, h->Xf, h->li };
	for (void *b : bufs) if (b) (void)hipFree(b);
	if (h->e0) (void)hipEventDestroy(h->e0);
	if (h->e1) (void)hipEventDestroy(h->e1);
	if (h->stream) (void)hipStreamDestroy(h->stream);
	delete h;
}

extern "C" int32_t nsh_mlp_param_count(void) { return (int32_t)P_N; }

extern "C" int nsh_mlp_get_params(nsh_mlp_t *h, float *params, float *grads)
{
	std::lock_guard<std::mutex> lk(h->mu);
	if (params && hipMemcpy(params, h->p, P_N * 4, hipMemcpyDeviceToHost) != hipSuccess) { nsh_set_error("mlp: copy failed"); return 0; }
	if (grads && hipMemcpy(grads, h->g, P_N * 4, hipMemcpyDeviceToHost) != hipSuccess) { nsh_set_error("mlp: copy failed"); return 0; }
	return 1;
}

static int ensure_train(nsh_mlp_t *h, int B)
{
	if (B <= h->cap) return 1;
	void *bufs[] = { h->X, h->H1, h->H2, h->dH1, h->dH2, h->Xt, h->t, h->y, h->dy, h->loss, h->lsum };
	for (void *b : bufs) if (b) (void)hipFree(b);
	h->X = h->H1 = h->H2 = h->dH1 = h->dH2 = nullptr;
	h->Xt = h->t = h->y = h->dy = h->loss = h->lsum = nullptr;
	h->cap = 0;
	MLPOK(hipMalloc((void**)&h->X, (size_t)B * IN * 2));
	MLPOK(hipMalloc((void**)&h->H1, (size_t)B * HID * 2));
	MLPOK(hipMalloc((void**)&h->H2, (size_t)B * HID * 2));
	MLPOK(hipMalloc((void**)&h->dH1, (size_t)B * HID * 2));
	MLPOK(hipMalloc((void**)&h->dH2, (size_t)B * HID * 2));
	MLPOK(hipMalloc((void**)&h->Xt, (size_t)B * IN * 4));
	MLPOK(hipMalloc((void**)&h->t, (size_t)B * 4));
	MLPOK(hipMalloc((void**)&h->y, (size_t)B * 4));
	MLPOK(hipMalloc((void**)&h->dy, (size_t)B * 4));
	MLPOK(hipMalloc((void**)&h->loss, (size_t)B * 4));
	MLPOK(hipMalloc((void**)&h->lsum, 4));
	h->cap = B;
	return 1;
fail:
	return 0;
}

// one step on a batch of B samples (B a multiple of 32): forward, backward into the gradient
// vector, and (when lr > 0) an Adam update. X: B x 64 FP32 features, t: B labels (0/1).
extern "C" int nsh_mlp_train_step(nsh_mlp_t *h, const float *X, const float *t, int32_t B, float lr, float *loss_out)
{
	if (!h || !X || !t || B < 32 || B % 32) { nsh_set_error("mlp: batch size %d must be a positive multiple of 32", B); return 0; }
	std::lock_guard<std::mutex> lk(h->mu);
	if (!ensure_train(h, B)) return 0;
	hipStream_t s = h->stream;
	float *xf = h->Xt;
	float lsum = 0.f;
	{
		MLPOK(hipMemcpyAsync(xf, X, (size_t)B * IN * 4, hipMemcpyHostToDevice, s));
		MLPOK(hipMemcpyAsync(h->t, t, (size_t)B * 4, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(((long)B * IN + 255) / 256), dim3(256), 0, s, xf, (long)B * IN,
			h->norm, h->norm + IN, h->X);
		hipLaunchKernelGGL(fill_f32_kernel, dim3((P_N + 255) / 256), dim3(256), 0, s, h->g, (long)P_N, 0.f);   // atomics below
		const int nrb = (B + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
		GemmArgs g = {};
		// H1 = relu(X W1 + b1): A = X [B][64], B(k,n) = W1[k][n] = W1T[n][k]
		g.A = h->X, g.sam = IN, g.sak = 1, g.B = h->W1Tb, g.sbn = IN, g.sbk = 1, g.M = B, g.N = HID, g.K = IN;
		g.epi = 1, g.Cb = h->H1, g.sbm = HID, g.bias = h->p + P_B1;
		gemm(s, g);
		// H2 = relu(H1 W2 + b2)
		g.A = h->H1, g.sam = HID, g.B = h->W2Tb, g.sbn = HID, g.K = HID, g.Cb = h->H2, g.bias = h->p + P_B2;
		gemm(s, g);
		hipLaunchKernelGGL(head_kernel, dim3((B + 3) / 4), dim3(256), 0, s, h->H2, h->p + P_W3, h->p + P_B3, h->t, B, h->y, h->dy, h->loss);
		hipLaunchKernelGGL(head_back_kernel, dim3(nrb), dim3(HID), 0, s, h->H2, h->dy, h->p + P_W3, B, h->g + P_W3, h->dH2);
		hipLaunchKernelGGL(vecsum_kernel, dim3(1), dim3(256), 0, s, h->dy, B, h->g + P_B3);
		// dW2 = H1^T dH2: A(m,k) = H1[k][m], B(k,n) = dH2[k][n]
		g = {};
		g.A = h->H1, g.sam = 1, g.sak = HID, g.B = h->dH2, g.sbn = 1, g.sbk = HID, g.M = HID, g.N = HID, g.K = B;
		g.epi = 0, g.Cf = h->g + P_W2, g.scm = HID;
		gemm(s, g);
		hipLaunchKernelGGL(colsum_kernel, dim3(nrb), dim3(HID), 0, s, h->dH2, B, HID, h->g + P_B2);
		// dH1 = (dH2 W2^T) [H1 > 0]: B(k,n) = W2[n][k] (row n of W2 has unit k stride)
		g = {};
		g.A = h->dH2, g.sam = HID, g.sak = 1, g.B = h->W2b, g.sbn = HID, g.sbk = 1, g.M = B, g.N = HID, g.K = HID;
		g.epi = 2, g.Cb = h->dH1, g.sbm = HID, g.mask = h->H1, g.smm = HID;
		gemm(s, g);
		// dW1 = X^T dH1
		g = {};
		g.A = h->X, g.sam = 1, g.sak = IN, g.B = h->dH1, g.sbn = 1, g.sbk = HID, g.M = IN, g.N = HID, g.K = B;
		g.epi = 0, g.Cf = h->g + P_W1, g.scm = HID;
		gemm(s, g);
		hipLaunchKernelGGL(colsum_kernel, dim3(nrb), dim3(HID), 0, s, h->dH1, B, HID, h->g + P_B1);
		hipLaunchKernelGGL(vecsum_kernel, dim3(1), dim3(256), 0, s, h->loss, B, h->lsum);
		if (lr > 0.f) {
			++h->step;
			AdamArgs a;
			a.p = h->p, a.g = h->g, a.m = h->m, a.v = h->v, a.n = (int)P_N;
			a.lr = lr, a.b1 = 0.9f, a.b2 = 0.999f, a.eps = 1e-8f;
			a.c1 = 1.f - std::pow(0.9f, (float)h->step), a.c2 = 1.f - std::pow(0.999f, (float)h->step);
			hipLaunchKernelGGL(adam_kernel, dim3((P_N + 255) / 256), dim3(256), 0, s, a);
			refresh_bf16(h);
		}
		MLPOK(hipGetLastError());
		MLPOK(hipMemcpyAsync(&lsum, h->lsum, 4, hipMemcpyDeviceToHost, s));
		MLPOK(hipStreamSynchronize(s));
		nsh_count_launch(12);
		if (loss_out) *loss_out = lsum / (float)B;
		return 1;
	}
fail:
	return 0;
}

// logits of n candidates (X: n x 64 FP32 features, host or device pointer)
extern "C" int nsh_mlp_infer(nsh_mlp_t *h, const float *X, int64_t n, int32_t x_on_device, float *logit)
{
	if (!h || n < 0 || (n > 0 && (!X || !logit))) { nsh_set_error("mlp: bad arguments"); return 0; }
	if (n == 0) return 1;
	std::lock_guard<std::mutex> lk(h->mu);
	hipStream_t s = h->stream;
	{
		if (n > h->icap) {
			void *bufs[] = { h->Xi, h->Xf, h->li };
			for (void *b : bufs) if (b) (void)hipFree(b);
			h->Xi = nullptr, h->Xf = h->li = nullptr, h->icap = 0;
			MLPOK(hipMalloc((void**)&h->Xi, (size_t)n * IN * 2));
			MLPOK(hipMalloc((void**)&h->Xf, (size_t)n * IN * 4));
			MLPOK(hipMalloc((void**)&h->li, (size_t)n * 4));
			h->icap = n;
		}
		const float *xd = X;
		if (!x_on_device) {
			MLPOK(hipMemcpyAsync(h->Xf, X, (size_t)n * IN * 4, hipMemcpyHostToDevice, s));
			xd = h->Xf;
		}
		MLPOK(hipEventRecord(h->e0, s));
		hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((n * IN + 255) / 256), dim3(256), 0, s, xd, n * IN, h->norm,
			h->norm + IN, h->Xi);
		hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, h->Xi, (int)n, h->W1Tb,
			h->p + P_B1, h->W2Tb, h->p + P_B2, h->p + P_W3, h->p + P_B3, h->li);
		MLPOK(hipEventRecord(h->e1, s));
		MLPOK(hipGetLastError());
		MLPOK(hipMemcpyAsync(logit, h->li, (size_t)n * 4, hipMemcpyDeviceToHost, s));
		MLPOK(hipStreamSynchronize(s));
		MLPOK(hipEventElapsedTime(&h->last_ms, h->e0, h->e1));
		nsh_count_launch(2);
		return 1;
	}
fail:
	return 0;
}

extern "C" float nsh_mlp_last_ms(nsh_mlp_t *h) { return h? h->last_ms : 0.f; }

// input standardisation applied on the device before the first layer: (x - mu) / sd per feature
extern "C" int nsh_mlp_set_norm(nsh_mlp_t *h, const float *mu, const float *sd)
{
	if (!h || !mu || !sd) { nsh_set_error("mlp: bad arguments"); return 0; }
	float nm[2 * IN];
	for (int f = 0; f < IN; ++f) {
		if (!(sd[f] > 0.f)) { nsh_set_error("mlp: sd[%d] must be positive", f); return 0; }
		nm[f] = mu[f], nm[IN + f] = 1.f / sd[f];
	}
	std::lock_guard<std::mutex> lk(h->mu);
	if (hipMemcpy(h->norm, nm, sizeof(nm), hipMemcpyHostToDevice) != hipSuccess) { nsh_set_error("mlp: copy failed"); return 0; }
	return 1;
}
