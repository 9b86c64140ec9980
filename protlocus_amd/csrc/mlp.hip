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
#include <type_traits>
#include <vector>
#include "nsh_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;
#define DEV static __device__ __forceinline__

constexpr int IN = 64, HID = 256;

// float -> bf16 bits, round to nearest even (finite inputs; gfx950 converts in hardware)
DEV u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }
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

// eight K-consecutive values of a row of an LDS tile whose rows are 16-byte aligned (one
// ds_read_b128, no alignment test, so the compiler can issue the loads of a K step ahead)
DEV bf16x8 lds8(const u16 *p, int stride, int row, int k)
{
	V8 x;
	x.q = *(const uint4*)(p + row * stride + k);
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
	// an operand with unit K stride, 16-byte aligned rows and a row stride of whole 16-byte units
	// is read with one vector load per fragment on every lane (a wave-uniform choice)
	const bool va = g.sak == 1 && !((uintptr_t)g.A & 15) && !(g.sam & 7);
	const bool vb = g.sbk == 1 && !((uintptr_t)g.B & 15) && !(g.sbn & 7);
	auto frag = [&](const u16 *p, long sr, long sk, bool vec, int row, int k) -> bf16x8 {
		if (vec) return *(const bf16x8*)(p + (long)row * sr + k);
		V8 x;
#pragma unroll
		for (int e = 0; e < 8; ++e) x.s[e] = p[(long)row * sr + (long)(k + e) * sk];
		return x.v;
	};
	// fragments of four K steps are loaded ahead of their products
	constexpr int D = 4;
	bf16x8 fa[D], fb[D];
#pragma unroll
	for (int d = 0; d < D; ++d)
		if (d * 16 < g.K) fa[d] = frag(g.A, g.sam, g.sak, va, m0 + r, d * 16 + kh), fb[d] = frag(g.B, g.sbn, g.sbk, vb, n0 + r, d * 16 + kh);
	f32x16 acc = {};
	for (int k0 = 0; k0 < g.K; k0 += 16 * D) {
#pragma unroll
		for (int d = 0; d < D; ++d) {
			if (k0 + d * 16 >= g.K) break;
			acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[d], fb[d], acc, 0, 0, 0);
			const int kn = k0 + (D + d) * 16;
			if (kn < g.K) fa[d] = frag(g.A, g.sam, g.sak, va, m0 + r, kn + kh), fb[d] = frag(g.B, g.sbn, g.sbk, vb, n0 + r, kn + kh);
		}
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

// the sum over lanes 0-31 in lane 31 and over lanes 32-63 in lane 63: running sums within each
// 16-lane row (DPP row_shr 1, 2, 4, 8), then row 0's total added to row 1 and row 2's to row 3
// (row_bcast15 on rows 1 and 3)
DEV float sum32(float v)
{
	auto sh = [](float x, auto ctrl) {
		return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), decltype(ctrl)::value, 0xf, 0xf, true));
	};
	v += sh(v, std::integral_constant<int, 0x111>());   // row_shr:1
	v += sh(v, std::integral_constant<int, 0x112>());   // row_shr:2
	v += sh(v, std::integral_constant<int, 0x114>());   // row_shr:4
	v += sh(v, std::integral_constant<int, 0x118>());   // row_shr:8
	v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xa, 0xf, true));   // row_bcast:15, rows 1 and 3
	return v;
}

// Fused inference, persistent: each work-group (8 waves, two per SIMD) takes blocks of 128 samples
// in turn. Wave w owns output columns [32w, 32w+32) of both hidden layers and keeps their B
// operands (its slices of W1T and W2T, 20 KB) in registers for the whole launch, so the weights are
// read once per wave. Per block: the raw FP32 features (loaded during the previous block) are
// standardised and rounded to BF16 into LDS; layers 1 and 2 run on MFMA with the activations in
// LDS (rows padded against bank conflicts, operands read with unconditional 16-byte loads) and
// four independent accumulators (the block's row tiles) in flight; layer 3 (256 -> 1) is folded
// into layer 2's epilogue as a dot product with w3, reduced over the wave's columns with DPP and
// over the waves through LDS. Measured on 1,048,576 samples: 0.38 ms, 455 TFLOP/s (a version with
// one wave per SIMD and two column tiles per wave: 0.46 ms; with shuffle reductions and an
// alignment test in every operand load: 0.87 ms).
constexpr int kInferRows = 128, kInferNT = 1, kInferWaves = HID / (32 * kInferNT);   // column tiles per wave, waves
constexpr int kInferBlocksPerCU = 1;   // work-groups per CU
__global__ void __launch_bounds__(64 * kInferWaves) mlp_infer_kernel(const float *X, int n, const float *mu,
		const float *isd, const u16 *W1T, const float *b1, const u16 *W2T, const float *b2, const float *w3,
		const float *b3, float *logit)
{
	constexpr int MT = kInferRows / 32, NT = 64 * kInferWaves;   // row tiles per block, threads
	constexpr int XV = kInferRows * IN / 4 / NT;                  // float4 of X per thread per block
	// rows padded by 16 bytes: the 32 rows an MFMA operand load reads start 4 banks apart
	constexpr int XS = IN + 8, HS = HID + 8;
	__shared__ __attribute__((aligned(16))) u16 xs[kInferRows * XS];     // 18 KB
	__shared__ __attribute__((aligned(16))) u16 h1[kInferRows * HS];     // 66 KB
	__shared__ float part[kInferWaves][kInferRows];
	const int L = threadIdx.x & 63, w = threadIdx.x >> 6, r = L & 31, kh = (L >> 5) * 8;
	bf16x8 wb1[kInferNT][IN / 16], wb2[kInferNT][HID / 16];
	float bias1[kInferNT], bias2[kInferNT], w3c[kInferNT];
#pragma unroll
	for (int nt = 0; nt < kInferNT; ++nt) {
		const int col = (w * kInferNT + nt) * 32 + r;
#pragma unroll
		for (int k = 0; k < IN / 16; ++k) wb1[nt][k] = load8(W1T, IN, 1, col, k * 16 + kh);
#pragma unroll
		for (int k = 0; k < HID / 16; ++k) wb2[nt][k] = load8(W2T, HID, 1, col, k * 16 + kh);
		bias1[nt] = b1[col], bias2[nt] = b2[col], w3c[nt] = w3[col];
	}
	const float bias3 = b3[0];
	const int nblk = (n + kInferRows - 1) / kInferRows;
	// X of a block: thread t holds float4 number t + NT * v; rows past n read as zero
	float4 xv[XV];
	auto load_x = [&](int blk) {
		const long s0 = (long)blk * kInferRows;
#pragma unroll
		for (int v = 0; v < XV; ++v) {
			const int t = threadIdx.x + NT * v, row = t / (IN / 4), c4 = (t % (IN / 4)) * 4;
			xv[v] = make_float4(0.f, 0.f, 0.f, 0.f);
			if (blk < nblk && s0 + row < n) xv[v] = *(const float4*)(X + (s0 + row) * IN + c4);
		}
	};
	load_x(blockIdx.x);
	for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
		const long s0 = (long)blk * kInferRows;
		__syncthreads();   // the previous block's readers of xs, h1 and part are done
		// stage X: (x - mu) * isd rounded to BF16 (rows past n: zero)
#pragma unroll
		for (int v = 0; v < XV; ++v) {
			const int t = threadIdx.x + NT * v, row = t / (IN / 4), c4 = (t % (IN / 4)) * 4;
			const bool in = s0 + row < n;
			const u16 o0 = in? f2bf((xv[v].x - mu[c4]) * isd[c4]) : 0, o1 = in? f2bf((xv[v].y - mu[c4 + 1]) * isd[c4 + 1]) : 0;
			const u16 o2 = in? f2bf((xv[v].z - mu[c4 + 2]) * isd[c4 + 2]) : 0, o3 = in? f2bf((xv[v].w - mu[c4 + 3]) * isd[c4 + 3]) : 0;
			*(uint2*)(xs + row * XS + c4) = make_uint2((unsigned)o0 | (unsigned)o1 << 16, (unsigned)o2 | (unsigned)o3 << 16);
		}
		load_x(blk + gridDim.x);   // the next block's features, in flight during this one
		__syncthreads();
		f32x16 acc[MT][kInferNT];
		// layer 1: this wave's columns for the block's rows
#pragma unroll
		for (int mt = 0; mt < MT; ++mt)
#pragma unroll
			for (int nt = 0; nt < kInferNT; ++nt) acc[mt][nt] = f32x16{};
#pragma unroll
		for (int k = 0; k < IN / 16; ++k)
#pragma unroll
			for (int mt = 0; mt < MT; ++mt) {
				const bf16x8 a = lds8(xs, XS, mt * 32 + r, k * 16 + kh);
#pragma unroll
				for (int nt = 0; nt < kInferNT; ++nt)
					acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wb1[nt][k], acc[mt][nt], 0, 0, 0);
			}
#pragma unroll
		for (int mt = 0; mt < MT; ++mt)
#pragma unroll
			for (int nt = 0; nt < kInferNT; ++nt)
#pragma unroll
				for (int i = 0; i < 16; ++i) {
					const int row = mt * 32 + 8 * (i >> 2) + 4 * (L >> 5) + (i & 3), col = (w * kInferNT + nt) * 32 + r;
					const float v = acc[mt][nt][i] + bias1[nt];
					h1[row * HS + col] = f2bf(v > 0.f? v : 0.f);
				}
		__syncthreads();
		// layer 2
#pragma unroll
		for (int mt = 0; mt < MT; ++mt)
#pragma unroll
			for (int nt = 0; nt < kInferNT; ++nt) acc[mt][nt] = f32x16{};
#pragma unroll
		for (int k = 0; k < HID / 16; ++k)
#pragma unroll
			for (int mt = 0; mt < MT; ++mt) {
				const bf16x8 a = lds8(h1, HS, mt * 32 + r, k * 16 + kh);
#pragma unroll
				for (int nt = 0; nt < kInferNT; ++nt)
					acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wb2[nt][k], acc[mt][nt], 0, 0, 0);
			}
		// layer 3: relu(h2) * w3 summed over the wave's columns
#pragma unroll
		for (int mt = 0; mt < MT; ++mt)
#pragma unroll
			for (int i = 0; i < 16; ++i) {
				float v = 0.f;
#pragma unroll
				for (int nt = 0; nt < kInferNT; ++nt) {
					const float h = acc[mt][nt][i] + bias2[nt];
					v += bf2f(f2bf(h > 0.f? h : 0.f)) * w3c[nt];   // the activation as stored in training
				}
				v = sum32(v);
				if (r == 31) part[w][mt * 32 + 8 * (i >> 2) + 4 * (L >> 5) + (i & 3)] = v;
			}
		__syncthreads();
		if (threadIdx.x < kInferRows && s0 + threadIdx.x < n) {
			float y = part[0][threadIdx.x];
#pragma unroll
			for (int v = 1; v < kInferWaves; ++v) y += part[v][threadIdx.x];
			logit[s0 + threadIdx.x] = y + bias3;
		}
	}
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
	float *Xf = nullptr, *li = nullptr;   // inference: features staged from the host, logits
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
		h->norm, h->Xt, h->t, h->y, h->dy, h->loss, h->lsum, h->Xf, h->li };
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
			void *bufs[] = { h->Xf, h->li };
			for (void *b : bufs) if (b) (void)hipFree(b);
			h->Xf = h->li = nullptr, h->icap = 0;
			MLPOK(hipMalloc((void**)&h->Xf, (size_t)n * IN * 4));
			MLPOK(hipMalloc((void**)&h->li, (size_t)n * 4));
			h->icap = n;
		}
		const float *xd = X;
		if (!x_on_device) {
			MLPOK(hipMemcpyAsync(h->Xf, X, (size_t)n * IN * 4, hipMemcpyHostToDevice, s));
			xd = h->Xf;
		} else if ((uintptr_t)X & 15) {   // the kernel reads rows as 16-byte vectors
			MLPOK(hipMemcpyAsync(h->Xf, X, (size_t)n * IN * 4, hipMemcpyDeviceToDevice, s));
			xd = h->Xf;
		}
		int dev = 0, ncu = 0;
		(void)hipGetDevice(&dev);
		(void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
		if (ncu <= 0) ncu = 256;
		const int64_t nblk = (n + kInferRows - 1) / kInferRows, grid = std::min<int64_t>(nblk, (int64_t)ncu * kInferBlocksPerCU);
		MLPOK(hipEventRecord(h->e0, s));
		hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)grid), dim3(64 * kInferWaves), 0, s, xd, (int)n, h->norm, h->norm + IN,
			h->W1Tb, h->p + P_B1, h->W2Tb, h->p + P_B2, h->p + P_W3, h->p + P_B3, h->li);
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
