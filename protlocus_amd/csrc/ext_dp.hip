// Batched spliced extension DP on gfx950 (score-only ns_global_gs16b with NS_F_EXT_LEFT/RIGHT;
// reference semantics: nasw-sse.c:360-459 and 596-741, recurrence nasw-sse.c:19-26).
//
// Execution model
//   * One wave64 solves one problem, or two problems of at most 256 padded columns (matrix
//     scores), one in each 16-bit half, which share the wave's fixed per-row work; waves are
//     persistent and pull work items from a queue ordered by decreasing cost, so long problems
//     start first and the tail is short.
//   * Row-synchronous: each lane owns C consecutive columns in each of two column halves
//     ([0, 64C) and [64C, 128C)); the two halves are packed into the 16-bit lanes of a
//     32-bit register and advanced together with VOP3P saturating instructions
//     (v_pk_add_i16/v_pk_sub_i16 clamp, v_pk_max_i16), i.e. two cells per instruction with
//     exactly the reference's int16 saturation.
//   * Within a row every cell except the horizontal insertion term depends only on previous
//     rows. Each lane runs the insertion chain over its own columns with a zero carry, then
//     one max-plus scan across lanes (DPP row_shr / row_bcast) yields every lane's exact
//     incoming insertion, which is folded into the lane's cells (max with a decaying carry).
//     Saturating subtraction is monotone, so this equals the reference's striped pass plus
//     lazy-F bit for bit.
//   * Per row the wave reduces the row maximum (padded columns included, columns beyond the
//     reference's padding excluded), applies the ie_coef*log2 length penalty from a host
//     table, updates the best row (snapshot of its H values in registers) and checks x-drop.
//   * Scores: the 22 x 22 matrix is staged once per work-group in LDS as a table of row
//     residue x column residue (a PSSM is staged per problem as an int8 profile of 22 residues
//     x 128C columns), so occupancy is bound by registers, not LDS. Per row one LDS read puts
//     the row residue's table row into a register (lane j = column residue j) and each column
//     takes its score from the lane of its residue with ds_bpermute, whose source addresses are
//     fixed per problem. The per-row residue, donor[i+1] and acceptor[i] come from a 4-byte
//     record per row prepared on the device; records and length penalties are read through the
//     scalar data cache a row before they are needed.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <mutex>
#include <vector>
#include "nsh_internal.h"
#include "ext_dp_types.h"
#include "global_dp_types.h"
#include "prep.h"

namespace nsh_ext {
void launch_group0(const ExtArgs &a, int grid, hipStream_t s);
void launch_group1(const ExtArgs &a, int grid, hipStream_t s);
void launch_group2(const ExtArgs &a, int grid, hipStream_t s);
void launch_group(int g, const ExtArgs &a, int grid, hipStream_t s)
{
	if (g == 0) launch_group0(a, grid, s);
	else if (g == 1) launch_group1(a, grid, s);
	else launch_group2(a, grid, s);
}
} // namespace nsh_ext

namespace {
using namespace nsh_ext;
using nsh_prep::PrepProb;
using nsh_prep::PrepConst;
constexpr int CMAX = kCmax;
constexpr int kPairMaxCols = 256;   // widest problem that shares a wave (4 columns per lane per half)
#define DEV static __device__ __forceinline__

__global__ void fill_i32_kernel(int32_t *p, int32_t n, int32_t v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------
// Device-side input preparation (nasw-sse.c:95-214): one work-group per problem writes
// rec[i] = nas[i] | donor[i+1] << 8 | acceptor[i] << 16 and the DP-order protein codes.
// ------------------------------------------------------------------------------------
DEV int wrap8(int v) { return (int)(int8_t)(uint8_t)(v & 0xff); }

// Every base a row's record needs lies within 8 positions of the row (codon i-2..i, donor up to
// i+8 on the left strand, acceptor down to i-6), so the work-group stages the nt4 codes of a chunk
// of rows plus that margin in LDS and reads them from there.
constexpr int kPrepChunk = 4096, kPrepHalo = 8;

struct PrepView {
	const uint8_t *ns, *ss; int nl, left, has_ss, max_spsc;
	const uint8_t *stage; int s0;          // nt4 codes of DP positions s0.. in LDS
	const int32_t *sp; int sp_null;
	__device__ int b(int x) const { return stage[x - s0]; } // nt4 code in DP order
	__device__ int spsc(int s) const { int v = (int)(int8_t)(s >> 1) - 64; return v > max_spsc? max_spsc : v; }
	__device__ int donor(int p) const {
		int d;
		if (!left) {
			if (p < nl - 3) {
				int t = 3;
				const int b1 = b(p + 1), b2 = b(p + 2);
				if (b1 == 2 && b2 == 3) t = (p + 3 < nl && (b(p + 3) == 0 || b(p + 3) == 2))? (b(p) == 2? -1 : 4) : 0;
				else if (b1 == 2 && b2 == 1 && b(p) == 2) t = 1;
				else if (b1 == 0 && b2 == 3) t = 2;
				d = wrap8(t < 0? 0 : sp[t]);
			} else d = wrap8(sp[3]);
			if (has_ss && p <= nl - 2) {
				const int s = ss[p + 1];
				if (s == 0xff) d = wrap8(d - sp_null);
				else if (!(s & 1)) d = wrap8(d - spsc(s));
			}
		} else {
			if (p < nl - 3) {
				int t = 3, pen_y = 0;
				const int b1 = b(p + 1), b2 = b(p + 2);
				if (b1 == 2 && b2 == 0) {
					t = (p + 3 < nl && (b(p + 3) == 1 || b(p + 3) == 3))? -1 : 0;
					for (int j = p + 5; j < nl && j < p + 8; ++j) { const int x = b(j); if (x != 1 && x != 3) pen_y += sp[5]; }
				} else if (b1 == 1 && b2 == 0) t = 2;
				d = wrap8(t < 0? 0 : sp[t]);
				if (t == -1 || t == 0) d = wrap8(d + pen_y);
			} else d = wrap8(sp[3]);
			if (has_ss && p <= nl - 1) {
				const int s = ss[nl - 1 - p];
				if (s == 0xff) d = wrap8(d - sp_null);
				else if (s & 1) d = wrap8(d - spsc(s));
			}
		}
		return d;
	}
	__device__ int acceptor(int p) const {
		int a;
		if (!left) {
			if (p >= 1 && p < nl) {
				int t = 3, pen_y = 0;
				const int bm1 = b(p - 1), b0 = b(p);
				if (bm1 == 0 && b0 == 2) {
					t = (p >= 2 && (b(p - 2) == 1 || b(p - 2) == 3))? -1 : 0;
					for (int j = p - 4; j >= 0 && j > p - 7; --j) { const int x = b(j); if (x != 1 && x != 3) pen_y += sp[5]; }
				} else if (bm1 == 0 && b0 == 1) t = 2;
				a = wrap8(t < 0? 0 : sp[t]);
				if (t == -1 || t == 0) a = wrap8(a + pen_y);
			} else a = wrap8(sp[3]);
			if (has_ss && p <= nl - 2) {
				const int s = ss[p + 1];
				if (s == 0xff) a = wrap8(a - sp_null);
				else if (s & 1) a = wrap8(a - spsc(s));
			}
		} else {
			if (p >= 1 && p < nl) {
				int t = 3;
				const int bm1 = b(p - 1), b0 = b(p);
				if (bm1 == 3 && b0 == 2) t = (p >= 2 && (b(p - 2) == 0 || b(p - 2) == 2))? ((p + 1 < nl && b(p + 1) == 2)? -1 : 4) : 0;
				else if (bm1 == 1 && b0 == 2 && p + 1 < nl && b(p + 1) == 1) t = 1;
				else if (bm1 == 3 && b0 == 0) t = 2;
				a = wrap8(t < 0? 0 : sp[t]);
			} else a = wrap8(sp[3]);
			if (has_ss && p <= nl - 1) {
				const int s = ss[nl - 1 - p];
				if (s == 0xff) a = wrap8(a - sp_null);
				else if (!(s & 1)) a = wrap8(a - spsc(s));
			}
		}
		return a;
	}
};

// idx (optional): the problems to prepare, one per work-group (subset runs)
__global__ void __launch_bounds__(256) ext_prep_kernel(const PrepProb *pp, const int32_t *idx, const uint8_t *raw_ns,
		const uint8_t *raw_ss, const uint8_t *raw_as, uint32_t *rec, uint8_t *aas, const PrepConst *pcg)
{
	__shared__ PrepConst pc;
	__shared__ uint8_t stage[kPrepChunk + 2 * kPrepHalo];
	const PrepProb P = pp[idx? idx[blockIdx.x] : (int32_t)blockIdx.x];
	for (int k = threadIdx.x; k < (int)sizeof(PrepConst); k += blockDim.x) ((uint8_t*)&pc)[k] = ((const uint8_t*)pcg)[k];
	__syncthreads();
	PrepView v;
	v.ns = (P.from_genome? pc.genome : raw_ns) + P.ns_off, v.ss = P.has_ss? raw_ss + P.ns_off : nullptr;
	v.nl = P.nl, v.left = P.left, v.has_ss = P.has_ss, v.max_spsc = (P.io + 1) / 2 - 1;
	v.stage = stage, v.sp = pc.sp, v.sp_null = pc.sp_null;
	const int X = pc.aa20['X'];
	const uint8_t *as = raw_as + P.as_off;
	for (int j = threadIdx.x; j < P.al; j += blockDim.x)
		aas[P.aa_off + j] = pc.aa20[as[P.left? P.al - 1 - j : j]];
	for (int i0 = 0; i0 < P.nl; i0 += kPrepChunk) {
		// the chunk's bases and the margin on both sides, as nt4 codes in DP order
		v.s0 = i0 - kPrepHalo;
		__syncthreads();
		for (int t = threadIdx.x; t < kPrepChunk + 2 * kPrepHalo; t += blockDim.x) {
			const int x = v.s0 + t;
			stage[t] = x >= 0 && x < P.nl? pc.nt4[v.ns[P.left? P.nl - 1 - x : x]] : 4;
		}
		__syncthreads();
		const int i1 = min(P.nl, i0 + kPrepChunk);
		for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
			// residue of the codon ending at DP position i: bases i-2, i-1, i in sequence order
			// (left: the reversed sequence read backwards, so DP positions i, i-1, i-2)
			int res = X;
			if (i >= 2 && (!P.left || i <= P.nl - 1)) {
				const int c0 = v.b(P.left? i : i - 2), c1 = v.b(i - 1), c2 = v.b(P.left? i - 2 : i);
				if (c0 < 4 && c1 < 4 && c2 < 4) res = pc.codon[c0 << 4 | c1 << 2 | c2];
			}
			const int d = v.donor(i + 1), ac = v.acceptor(i);
			rec[P.rec_off + i] = (uint32_t)res | ((uint32_t)(d & 0xff) << 8) | ((uint32_t)(ac & 0xff) << 16);
		}
	}
}

} // namespace

void nsh_prep::launch(const PrepProb *pp, int32_t n, const uint8_t *raw_ns, const uint8_t *raw_ss, const uint8_t *raw_as,
		uint32_t *rec, uint8_t *aas, const PrepConst *pc, hipStream_t s)
{
	hipLaunchKernelGGL(ext_prep_kernel, dim3(n), dim3(256), 0, s, pp, (const int32_t*)nullptr, raw_ns, raw_ss, raw_as, rec, aas, pc);
}

// ======================================================================================
// host side: plans
// ======================================================================================
#define HIPOK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { nsh_set_error("%s: %s", #x, hipGetErrorString(e_)); goto fail; } } while (0)

// A plan's device buffers, pinned staging, streams and events live in a shell. Destroying a plan
// returns its shell to a small pool (at most kPoolShells, each at most kPoolBytes of device
// buffers), and a new plan takes a pooled shell and only grows the buffers it needs larger, so
// repeated extension batches of similar size allocate nothing.
template <typename T>
struct DBuf {
	T *p = nullptr;
	size_t cap = 0;                       // elements
	hipError_t ensure(size_t n)
	{
		if (n <= cap && p) return hipSuccess;
		if (p) (void)hipFree(p);
		p = nullptr, cap = 0;
		const size_t want = std::max(n + (n >> 3), (size_t)64);
		const hipError_t e = hipMalloc((void**)&p, want * sizeof(T) + 16);
		if (e == hipSuccess) cap = want;
		return e;
	}
	size_t bytes() const { return p? cap * sizeof(T) : 0; }
	void release() { if (p) (void)hipFree(p); p = nullptr, cap = 0; }
};

struct nsh_ext_plan_s {
	int32_t n = 0;
	bool from_genome = false;              // some problems read their bases from a genome in HBM
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
	hipEvent_t evf = nullptr;              // fork point of the group streams (no timing: usable in a capture)
	hipStream_t gs[kGroups] = {};          // one stream per column-count group
	// graph mode: the whole run (preparation, group kernels, wide problems) captured once and replayed
	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	int32_t ggrid[kGroups] = {};           // the persistent grids the graph was captured with
	hipEvent_t gev[kGroups] = {};          // completion of each group's launch
	// device buffers (grow-only in a pooled shell)
	DBuf<uint8_t> b_ns, b_ss, b_as, b_aas;
	DBuf<uint32_t> b_rec;
	DBuf<int8_t> b_mat, b_pssm;
	DBuf<int32_t> b_pen, b_counter, b_out, b_idx_sub;
	DBuf<ExtProb> b_prob, b_prob_sub;
	DBuf<PrepProb> b_prep;
	DBuf<PrepConst> b_pc;
	DBuf<float> b_feat;
	DBuf<nsh_glob::GlobProb> b_wide, b_wide_sub;
	// pinned host staging of the inputs (grow-only)
	void *h_stage = nullptr;
	size_t h_stage_cap = 0;
	// the current plan's view of the buffers
	uint8_t *d_ns = nullptr, *d_ss = nullptr, *d_as = nullptr, *d_aas = nullptr;
	uint32_t *d_rec = nullptr;
	int8_t *d_mat = nullptr, *d_pssm = nullptr;
	int32_t *d_pen = nullptr, *d_counter = nullptr, *d_out = nullptr;
	ExtProb *d_prob = nullptr;
	PrepProb *d_prep = nullptr;
	PrepConst *d_pc = nullptr;
	ExtArgs args[kGroups] = {};
	int32_t grid[kGroups] = {};
	int32_t wpc[kGroups] = {};   // persistent waves per CU of each group (defaults, or the autotuner's)
	double cells_window = 0;     // sum nl*al
	std::vector<int32_t> al;     // per input index
	std::vector<int32_t> out_h;  // 4 per problem
	int ran = 0;
	int ncu = 256;
	int32_t max_al = 0, max_nl = 0;
	std::vector<ExtProb> prob_h;             // descriptors in queue order (host copy)
	std::vector<ExtProb> prob_in;            // the narrow problems' descriptors in input order
	bool pssm_plan = false;                  // some problem scores with a PSSM (no pairs then)
	int32_t pair_rows = INT32_MAX;           // problems of more rows run one per wave (the autotuner's choice)
	bool merged = false;                     // group 0 is the merged C <= 8 queue (launch_group01)
	int32_t grp_first[kGroups] = {}, grp_n[kGroups] = {};
	// problems wider than 128 * kCmax padded columns: the work-group-per-problem kernel in extension
	// mode (global_dp_impl.h), in buckets of equal shape, on their own stream
	std::vector<nsh_glob::GlobProb> wide_h;
	std::vector<nsh_glob::GlobProb> wide_sub_h;
	nsh_glob::GlobProb *d_wide = nullptr, *d_wide_sub = nullptr;
	hipStream_t ws = nullptr;
	hipEvent_t wev = nullptr;
	nsh_glob::GlobArgs wargs = {};
	// subset runs and features
	ExtProb *d_prob_sub = nullptr;
	std::vector<ExtProb> sub_h;
	int32_t *d_idx_sub = nullptr;            // kept input indices (their inputs are prepared)
	std::vector<int32_t> idx_h;
	float *d_feat = nullptr;

	size_t device_bytes() const
	{
		return b_ns.bytes() + b_ss.bytes() + b_as.bytes() + b_aas.bytes() + b_rec.bytes() + b_mat.bytes() + b_pssm.bytes() +
			b_pen.bytes() + b_counter.bytes() + b_out.bytes() + b_idx_sub.bytes() + b_prob.bytes() + b_prob_sub.bytes() +
			b_prep.bytes() + b_pc.bytes() + b_feat.bytes() + b_wide.bytes() + b_wide_sub.bytes();
	}
	void drop_graph()
	{
		if (gexec) (void)hipGraphExecDestroy(gexec);
		if (graph) (void)hipGraphDestroy(graph);
		gexec = nullptr, graph = nullptr;
	}
	void reset()   // forget the last plan, keep the resources
	{
		drop_graph();
		n = 0, ran = 0, cells_window = 0, max_al = max_nl = 0;
		al.clear(), out_h.clear(), prob_h.clear(), wide_h.clear(), wide_sub_h.clear(), sub_h.clear(), idx_h.clear();
		prob_in.clear(), pssm_plan = false, pair_rows = INT32_MAX, merged = false;
		d_ns = d_ss = d_as = d_aas = nullptr, d_rec = nullptr, d_mat = d_pssm = nullptr;
		d_pen = d_counter = d_out = nullptr, d_prob = nullptr, d_prep = nullptr, d_pc = nullptr;
		d_wide = d_wide_sub = nullptr, d_prob_sub = nullptr, d_idx_sub = nullptr, d_feat = nullptr;
		for (int g = 0; g < kGroups; ++g) args[g] = ExtArgs{}, grid[g] = wpc[g] = grp_first[g] = grp_n[g] = 0;
		wargs = nsh_glob::GlobArgs{};
	}
	void release_all()
	{
		b_ns.release(), b_ss.release(), b_as.release(), b_aas.release(), b_rec.release(), b_mat.release(), b_pssm.release();
		b_pen.release(), b_counter.release(), b_out.release(), b_idx_sub.release(), b_prob.release(), b_prob_sub.release();
		b_prep.release(), b_pc.release(), b_feat.release(), b_wide.release(), b_wide_sub.release();
		if (h_stage) (void)hipHostFree(h_stage);
		h_stage = nullptr, h_stage_cap = 0;
		if (ev0) (void)hipEventDestroy(ev0);
		if (ev1) (void)hipEventDestroy(ev1);
		if (ev2) (void)hipEventDestroy(ev2);
		if (evf) (void)hipEventDestroy(evf);
		for (int g = 0; g < kGroups; ++g) {
			if (gev[g]) (void)hipEventDestroy(gev[g]);
			if (gs[g]) (void)hipStreamDestroy(gs[g]);
		}
		if (wev) (void)hipEventDestroy(wev);
		if (ws) (void)hipStreamDestroy(ws);
		if (stream) (void)hipStreamDestroy(stream);
		ev0 = ev1 = ev2 = evf = wev = nullptr, ws = stream = nullptr;
		for (int g = 0; g < kGroups; ++g) gev[g] = nullptr, gs[g] = nullptr;
	}
};

namespace {
constexpr int kPoolShells = 4;
constexpr size_t kPoolBytes = (size_t)2 << 30;
std::mutex g_pool_mu;
std::vector<nsh_ext_plan_t*> g_pool;

// a shell with its streams and events (pooled or new); nullptr on failure
nsh_ext_plan_t *shell_acquire()
{
	{
		std::lock_guard<std::mutex> lk(g_pool_mu);
		if (!g_pool.empty()) { nsh_ext_plan_t *p = g_pool.back(); g_pool.pop_back(); return p; }
	}
	nsh_ext_plan_t *p = new nsh_ext_plan_s();
	bool ok = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess &&
		hipEventCreate(&p->ev0) == hipSuccess && hipEventCreate(&p->ev1) == hipSuccess && hipEventCreate(&p->ev2) == hipSuccess &&
		hipEventCreateWithFlags(&p->evf, hipEventDisableTiming) == hipSuccess &&
		hipStreamCreateWithFlags(&p->ws, hipStreamNonBlocking) == hipSuccess &&
		hipEventCreateWithFlags(&p->wev, hipEventDisableTiming) == hipSuccess;
	for (int g = 0; ok && g < kGroups; ++g)
		ok = hipStreamCreateWithFlags(&p->gs[g], hipStreamNonBlocking) == hipSuccess &&
			hipEventCreateWithFlags(&p->gev[g], hipEventDisableTiming) == hipSuccess;
	if (!ok) { nsh_set_error("extension plan: stream set-up failed"); p->release_all(); delete p; return nullptr; }
	return p;
}

void shell_release(nsh_ext_plan_t *p)
{
	if (p->stream) (void)hipStreamSynchronize(p->stream);
	p->reset();
	{
		std::lock_guard<std::mutex> lk(g_pool_mu);
		if ((int)g_pool.size() < kPoolShells && p->device_bytes() <= kPoolBytes) { g_pool.push_back(p); return; }
	}
	p->release_all();
	delete p;
}
} // namespace

// The persistent queue of a plan's narrow problems (p->prob_in) into p->prob_h: work items in
// queue order, then the pair partners. Two problems per wave (matrix scores only): problems of at
// most 256 padded columns, at least 3 rows and at most pair_rows rows, paired with a problem of the
// same column count per lane (ceil(npad / 64)) and the nearest length; the first of a pair is the
// queue item. Items are ordered small-C group first (its launch), then by the most wave time first
// in each group (rows x per-row instructions, about 41 + 30 C for one problem and 49 + 30 C for a
// pair). n_grp receives the items per column-count group.
static int plan_group(const nsh_ext_plan_t *p, int32_t C) { return p->merged && C <= 8? 0 : group_of(C); }

static void plan_order(nsh_ext_plan_t *p, int32_t pair_rows, int32_t n_grp[kGroups])
{
	const std::vector<ExtProb> &prob = p->prob_in;
	const int32_t n = (int32_t)prob.size();
	std::vector<int32_t> partner(n, -1), order;
	std::vector<char> second(n, 0);
	for (int g = 0; g < kGroups; ++g) n_grp[g] = 0;
	if (!p->pssm_plan) {
		std::vector<int32_t> cand;
		for (int32_t k = 0; k < n; ++k)
			if (prob[k].npad <= kPairMaxCols && prob[k].nl >= 3 && prob[k].nl <= pair_rows) cand.push_back(k);
		auto cp = [&](int32_t k) { return (prob[k].npad + 63) / 64; };
		std::stable_sort(cand.begin(), cand.end(), [&](int32_t x, int32_t y) {
			return cp(x) != cp(y)? cp(x) < cp(y) : prob[x].nl > prob[y].nl; });
		for (size_t t = 0; t + 1 < cand.size();) {
			if (cp(cand[t]) != cp(cand[t + 1])) { ++t; continue; }
			partner[cand[t]] = cand[t + 1], second[cand[t + 1]] = 1;
			t += 2;
		}
	}
	auto cost = [&](int32_t k) -> int64_t {
		if (partner[k] < 0) return (int64_t)prob[k].nl * (41 + 30 * prob[k].C);
		return (int64_t)std::max(prob[k].nl, prob[partner[k]].nl) * (49 + 30 * ((prob[k].npad + 63) / 64));
	};
	for (int32_t k = 0; k < n; ++k) if (!second[k]) order.push_back(k), ++n_grp[plan_group(p, prob[k].C)];
	std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
		const int bx = plan_group(p, prob[x].C), by = plan_group(p, prob[y].C);
		if (bx != by) return bx < by;
		return cost(x) > cost(y); });
	const int32_t n_items = (int32_t)order.size();
	for (int32_t k = 0; k < n_items; ++k) if (partner[order[k]] >= 0) order.push_back(partner[order[k]]);
	p->prob_h.resize(order.size());
	for (size_t k = 0; k < order.size(); ++k) p->prob_h[k] = prob[order[k]];
	for (int32_t k = 0, t = n_items; k < n_items; ++k) {
		if (partner[order[k]] < 0) continue;
		ExtProb &e = p->prob_h[k];
		e.pair = t++, e.C = (e.npad + 63) / 64;   // the partner's index; C per lane per problem
	}
}

// the groups' queue views of p->prob_h after plan_order (descriptors already on the device)
static void plan_set_groups(nsh_ext_plan_t *p, const int32_t n_grp[kGroups])
{
	int64_t first = 0;
	for (int g = 0; g < kGroups; ++g) {
		ExtArgs &a = p->args[g];
		p->grp_first[g] = (int32_t)first, p->grp_n[g] = n_grp[g];
		a.prob = p->d_prob + first, a.n = n_grp[g], a.pbase = p->d_prob;
		first += n_grp[g];
		p->grid[g] = (int32_t)std::min<int64_t>(a.n, (int64_t)p->ncu * p->wpc[g]);
	}
}

// the queue rebuilt with another pairing limit and uploaded (the plan's stream is idle after this)
static int plan_requeue(nsh_ext_plan_t *p, int32_t pair_rows)
{
	int32_t n_grp[kGroups];
	HIPOK(hipStreamSynchronize(p->stream));
	p->drop_graph();   // a captured graph holds the old queue sizes
	p->pair_rows = pair_rows;
	plan_order(p, pair_rows, n_grp);
	if (!p->prob_h.empty())
		HIPOK(hipMemcpyAsync(p->d_prob, p->prob_h.data(), p->prob_h.size() * sizeof(ExtProb), hipMemcpyHostToDevice, p->stream));
	HIPOK(hipStreamSynchronize(p->stream));
	plan_set_groups(p, n_grp);
	return 1;
fail:
	return 0;
}

int nsh_ext_req_fits(const nsh_ext_req &q, int32_t k)
{
	// npad is the protein rounded up to a stripe width of 8 or 16; the kernels rely on a multiple of 8
	if (q.al < 1 || q.nl < 0 || q.npad < q.al || q.npad % 8) { nsh_set_error("request %d: nl=%d al=%d npad=%d", k, q.nl, q.al, q.npad); return 0; }
	if (q.npad > 64 * nsh_glob::kCMaxG * nsh_glob::kNWMax) {
		nsh_set_error("request %d: padded protein length %d exceeds %d", k, q.npad, 64 * nsh_glob::kCMaxG * nsh_glob::kNWMax);
		return 0;
	}
	return 1;
}

nsh_ext_plan_t *nsh_ext_plan_build(int32_t n, const nsh_ext_req *req, const ns_opt_t *opt)
{
	if (!nsh_ensure_device()) return nullptr;
	if (n < 1) { nsh_set_error("empty extension plan"); return nullptr; }
	const int32_t asize = opt->asize;
	if (asize != 22) { nsh_set_error("asize %d not supported (22 expected)", asize); return nullptr; }
	for (int t = 0; t < 64; ++t)
		if (opt->codon[t] >= 22) { nsh_set_error("codon table entry %d out of range", t); return nullptr; }
	for (int t = 0; t < 256; ++t)
		if (opt->aa20[t] >= 22 || opt->nt4[t] > 4) { nsh_set_error("alphabet table entry %d out of range", t); return nullptr; }
	if (opt->ge < 0 || opt->go < 0 || (int64_t)opt->ge * CMAX * 64 > 32767) {
		nsh_set_error("gap penalties go=%d ge=%d out of the supported range", opt->go, opt->ge); return nullptr;
	}
	nsh_step_clock clk("nsh_ext_plan_build");
	int64_t l_ns = 0, l_raw = 0, l_as = 0, l_pssm = 0;
	const uint8_t *genome = nullptr;
	int32_t max_nl = 2, any_ss = 0, any_long = 0, n_grp[kGroups] = {};
	std::vector<ExtProb> prob(n);
	std::vector<PrepProb> prep(n);
	std::vector<int32_t> order;
	std::vector<char> is_wide(n, 0);
	for (int32_t k = 0; k < n; ++k) {
		const nsh_ext_req &q = req[k];
		int32_t C = (q.npad + 127) / 128;
		if (!nsh_ext_req_fits(q, k)) return nullptr;
		if (C < 1) C = 1;
		is_wide[k] = C > CMAX;
		ExtProb &e = prob[k];
		e.nl = q.nl, e.al = q.al, e.npad = q.npad, e.io = q.io, e.out = k, e.C = C, e.left = q.left, e.mode = q.mode;
		e.rec_off = l_ns, e.aa_off = l_as, e.pssm_off = -1;
		PrepProb &r = prep[k];
		r = PrepProb{};
		if (q.g_off >= 0) {   // bases read from the oriented genome in HBM
			if (!q.d_g || q.ss || (genome && genome != q.d_g)) { nsh_set_error("request %d: bad genome window", k); return nullptr; }
			genome = q.d_g, r.from_genome = 1, r.ns_off = q.g_off;
		} else r.ns_off = l_raw, l_raw += q.nl + 1;
		r.as_off = l_as, r.rec_off = l_ns, r.aa_off = l_as;
		r.nl = q.nl, r.al = q.al, r.left = q.left, r.io = q.io, r.has_ss = q.ss != nullptr;
		if (q.pssm) e.pssm_off = l_pssm, l_pssm += (int64_t)q.al * asize;
		l_ns += q.nl + 1, l_as += q.al + 1;
		if (q.nl > max_nl) max_nl = q.nl;
		any_ss |= q.ss != nullptr;
		if (!is_wide[k]) ++n_grp[group_of(C)];
		any_long |= (int64_t)q.nl * C >= kLongWork;
	}
	nsh_ext_plan_t *p = shell_acquire();
	if (!p) return nullptr;
	p->n = n;
	p->from_genome = genome != nullptr;
	p->al.resize(n);
	for (int32_t k = 0; k < n; ++k) {
		p->max_nl = std::max(p->max_nl, req[k].nl), p->max_al = std::max(p->max_al, req[k].al);
		p->al[k] = req[k].al;
		p->cells_window += (double)req[k].nl * req[k].al;
	}
	// wide problems: descriptors of the work-group kernel, in buckets of equal shape
	for (int32_t k = 0; k < n; ++k) {
		if (!is_wide[k]) continue;
		const ExtProb &e = prob[k];
		nsh_glob::GlobProb g = {};
		g.rec_off = e.rec_off, g.aa_off = e.aa_off, g.tr_off = 0, g.pssm_off = e.pssm_off;
		g.nl = e.nl, g.al = e.al, g.io = e.io, g.out = e.out, g.left = e.left, g.npad = e.npad, g.mode = e.mode;
		g.slen = 1;
		nsh_glob::shape(e.npad, g.C, g.NW);
		p->wide_h.push_back(g);
	}
	std::stable_sort(p->wide_h.begin(), p->wide_h.end(), [](const nsh_glob::GlobProb &x, const nsh_glob::GlobProb &y) {
		return x.NW != y.NW? x.NW > y.NW : x.C > y.C; });
	for (int32_t k = 0; k < n; ++k) if (!is_wide[k]) p->prob_in.push_back(prob[k]);
	p->pssm_plan = l_pssm > 0;
	p->merged = !p->pssm_plan && any_long && n_grp[0] > 0 && n_grp[1] > 0;
	clk.lap("  problems", n);
	plan_order(p, p->pair_rows, n_grp);
	clk.lap("  queue order", n);
	const int32_t n_narrow = (int32_t)p->prob_h.size();
	p->out_h.assign((size_t)4 * n, 0);
	// inputs staged in pinned memory: ns | as | ss | pssm | pen | prob | prep | pc | mat | wide
	const int32_t n_pen = max_nl + 2;
	const size_t s_ns = 0, s_as = (s_ns + l_raw + 255) & ~(size_t)255, s_ss = (s_as + l_as + 255) & ~(size_t)255;
	const size_t s_pssm = (s_ss + (any_ss? l_raw : 0) + 255) & ~(size_t)255;
	const size_t s_pen = (s_pssm + l_pssm + 255) & ~(size_t)255;
	const size_t s_prob = (s_pen + (size_t)n_pen * 4 + 255) & ~(size_t)255;
	const size_t s_prep = (s_prob + (size_t)n_narrow * sizeof(ExtProb) + 255) & ~(size_t)255;
	const size_t s_pc = (s_prep + (size_t)n * sizeof(PrepProb) + 255) & ~(size_t)255;
	const size_t s_mat = (s_pc + sizeof(PrepConst) + 255) & ~(size_t)255;
	const size_t s_wide = (s_mat + 484 + 255) & ~(size_t)255;
	const size_t s_end = s_wide + p->wide_h.size() * sizeof(nsh_glob::GlobProb);
	uint8_t *hs = nullptr;
	if (p->h_stage_cap < s_end) {
		if (p->h_stage) (void)hipHostFree(p->h_stage);
		p->h_stage = nullptr, p->h_stage_cap = 0;
		const size_t want = std::max(s_end + (s_end >> 3), (size_t)1 << 20);
		HIPOK(hipHostMalloc(&p->h_stage, want, hipHostMallocDefault));
		p->h_stage_cap = want;
	}
	hs = (uint8_t*)p->h_stage;
	clk.lap("  pinned staging", (int64_t)s_end);
	// the strings into the pinned block on the host threads, 1,024 problems at a time
	nsh_parallel_for(((int64_t)n + 1023) / 1024, n >= 4096? 16 : 1, [&](int64_t b) {
		for (int32_t k = (int32_t)(b * 1024); k < std::min<int64_t>(n, (b + 1) * 1024); ++k) {
			const nsh_ext_req &q = req[k];
			if (q.nl && !prep[k].from_genome) std::memcpy(hs + s_ns + prep[k].ns_off, q.ns, q.nl);
			std::memcpy(hs + s_as + prep[k].as_off, q.as, q.al);
			if (any_ss && !prep[k].from_genome) {
				if (q.ss) std::memcpy(hs + s_ss + prep[k].ns_off, q.ss, q.nl);
				else std::memset(hs + s_ss + prep[k].ns_off, 0xff, q.nl);
			}
			if (q.pssm) std::memcpy(hs + s_pssm + prob[k].pssm_off, q.pssm, (size_t)q.al * asize);
		}
	});
	{
		int32_t *pen = (int32_t*)(hs + s_pen);
		for (int32_t d = 0; d < n_pen; ++d) pen[d] = nsh_ie_penalty(opt->ie_coef, d);
		PrepConst &pc = *(PrepConst*)(hs + s_pc);
		std::memcpy(pc.nt4, opt->nt4, 256), std::memcpy(pc.aa20, opt->aa20, 256), std::memcpy(pc.codon, opt->codon, 64);
		for (int t = 0; t < 6; ++t) pc.sp[t] = opt->sp[t];
		pc.sp_null = opt->sp_null_bonus;
		pc.genome = genome;
		if (n_narrow) std::memcpy(hs + s_prob, p->prob_h.data(), (size_t)n_narrow * sizeof(ExtProb));
		std::memcpy(hs + s_prep, prep.data(), (size_t)n * sizeof(PrepProb));
		std::memcpy(hs + s_mat, opt->sc, (size_t)asize * asize);
		if (!p->wide_h.empty()) std::memcpy(hs + s_wide, p->wide_h.data(), p->wide_h.size() * sizeof(nsh_glob::GlobProb));
	}
	clk.lap("  pack", n);
	HIPOK(p->b_ns.ensure(std::max<int64_t>(l_raw, 1))); HIPOK(p->b_as.ensure(l_as)); HIPOK(p->b_aas.ensure(l_as));
	HIPOK(p->b_rec.ensure(l_ns));
	if (any_ss) HIPOK(p->b_ss.ensure(std::max<int64_t>(l_raw, 1)));
	HIPOK(p->b_mat.ensure((size_t)asize * asize));
	if (l_pssm) HIPOK(p->b_pssm.ensure(l_pssm));
	HIPOK(p->b_pen.ensure(n_pen)); HIPOK(p->b_counter.ensure(kGroups)); HIPOK(p->b_out.ensure((size_t)4 * n));
	HIPOK(p->b_prob.ensure(std::max(1, n_narrow))); HIPOK(p->b_prep.ensure(n)); HIPOK(p->b_pc.ensure(1));
	if (!p->wide_h.empty()) HIPOK(p->b_wide.ensure(p->wide_h.size()));
	p->d_ns = p->b_ns.p, p->d_as = p->b_as.p, p->d_aas = p->b_aas.p, p->d_rec = p->b_rec.p;
	p->d_ss = any_ss? p->b_ss.p : nullptr, p->d_mat = p->b_mat.p, p->d_pssm = l_pssm? p->b_pssm.p : nullptr;
	p->d_pen = p->b_pen.p, p->d_counter = p->b_counter.p, p->d_out = p->b_out.p, p->d_prob = p->b_prob.p;
	p->d_prep = p->b_prep.p, p->d_pc = p->b_pc.p, p->d_wide = p->wide_h.empty()? nullptr : p->b_wide.p;
	{   // from pinned memory: asynchronous; the stream orders them before the runs
		auto cp = [&](void *dst, size_t off, size_t bytes) {
			return bytes == 0 || hipMemcpyAsync(dst, hs + off, bytes, hipMemcpyHostToDevice, p->stream) == hipSuccess;
		};
		const bool ok = cp(p->d_ns, s_ns, l_raw) && cp(p->d_as, s_as, l_as) && (!any_ss || cp(p->d_ss, s_ss, l_raw)) &&
			(!l_pssm || cp(p->d_pssm, s_pssm, l_pssm)) && cp(p->d_pen, s_pen, (size_t)n_pen * 4) &&
			cp(p->d_prob, s_prob, (size_t)n_narrow * sizeof(ExtProb)) && cp(p->d_prep, s_prep, (size_t)n * sizeof(PrepProb)) &&
			cp(p->d_pc, s_pc, sizeof(PrepConst)) && cp(p->d_mat, s_mat, (size_t)asize * asize) &&
			(p->wide_h.empty() || cp(p->d_wide, s_wide, p->wide_h.size() * sizeof(nsh_glob::GlobProb)));
		if (!ok) { nsh_set_error("extension plan: copy to the device failed"); goto fail; }
	}
	if (clk.on) {
		(void)hipStreamSynchronize(p->stream);
		clk.lap("  device buffers, upload", (int64_t)s_end);
	}
	{
		int dev = 0, ncu = 0;
		(void)hipGetDevice(&dev);
		(void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
		if (ncu <= 0) ncu = 256;
		p->ncu = ncu;
		int64_t first = 0;
		for (int g = 0; g < kGroups; ++g) {
			ExtArgs &a = p->args[g];
			p->grp_first[g] = (int32_t)first, p->grp_n[g] = n_grp[g];
			a.prob = p->d_prob + first, a.n = n_grp[g], a.pbase = p->d_prob;
			first += n_grp[g];
			a.rec = p->d_rec, a.aas = p->d_aas, a.pssm = p->d_pssm, a.mat = p->d_mat, a.asize = asize;
			a.pen = p->d_pen, a.pen_n = n_pen;
			a.go = opt->go, a.ge = opt->ge, a.fs = opt->fs, a.xdrop = opt->xdrop, a.eb = opt->end_bonus;
			a.counter = p->d_counter + g, a.out = p->d_out;
			// persistent waves: about as many as the CUs keep resident
			p->wpc[g] = l_pssm? kGroupWavesPerCUPssm[g] : p->merged && g == 0? kMergedWavesPerCU : kGroupWavesPerCU[g];
			p->grid[g] = (int32_t)std::min<int64_t>(a.n, (int64_t)ncu * p->wpc[g]);
		}
		nsh_glob::GlobArgs &w = p->wargs;
		w.rec = p->d_rec, w.aas = p->d_aas, w.pssm = p->d_pssm, w.mat = p->d_mat, w.asize = asize;
		w.go = opt->go, w.ge = opt->ge, w.fs = opt->fs, w.trace = nullptr, w.score = nullptr;
		w.pen = p->d_pen, w.pen_n = n_pen, w.xdrop = opt->xdrop, w.eb = opt->end_bonus, w.ext_out = p->d_out;
	}
	return p;
fail:
	shell_release(p);
	return nullptr;
}

extern "C" nsh_ext_plan_t *nsh_ext_plan_create(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl,
		const char *as, const int64_t *as_off, const int32_t *al, const uint8_t *ss,
		const uint8_t *dirs, const ns_opt_t *opt)
{
	if (n < 1 || !ns || !as || !nl || !al || !opt) { nsh_set_error("nsh_ext_plan_create: bad arguments"); return nullptr; }
	std::vector<nsh_ext_req> req(n);
	// no direction at all: plain score-only global DP (ns_global_gs16b without CIGAR or EXT flags)
	const bool global = !dirs && !(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	for (int32_t k = 0; k < n; ++k) {
		nsh_ext_req &q = req[k];
		const int32_t flag = dirs? (dirs[k]? NS_F_EXT_LEFT : NS_F_EXT_RIGHT) : (opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
		q.ns = ns + (ns_off? ns_off[k] : 0), q.nl = nl[k];
		q.as = as + (as_off? as_off[k] : 0), q.al = al[k];
		q.ss = ss? ss + (ns_off? ns_off[k] : 0) : nullptr;
		q.left = (flag & NS_F_EXT_LEFT) != 0;
		q.io = opt->io;
		q.mode = global;
		q.pssm = opt->pssm;
		const int32_t w = nsh_host_stripe_width(global? 0 : flag, q.al);
		q.npad = (q.al + w - 1) / w * w;
	}
	return nsh_ext_plan_build(n, req.data(), opt);
}

// results of problems a subset run skips: nt_len = aa_len = rows = 0, score = INT32_MIN
__global__ void clear_out_kernel(int32_t *out, int32_t n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[4 * i] = 0, out[4 * i + 1] = 0, out[4 * i + 2] = INT_MIN, out[4 * i + 3] = 0;
}

// the wide problems' buckets (runs of equal NW and C in list) on the plan's wide stream
static int launch_wide(nsh_ext_plan_t *p, const std::vector<nsh_glob::GlobProb> &list, const nsh_glob::GlobProb *d_list)
{
	if (list.empty()) return 1;
	HIPOK(hipStreamWaitEvent(p->ws, p->evf, 0));
	for (size_t k = 0; k < list.size();) {
		size_t e = k;
		while (e < list.size() && list[e].NW == list[k].NW && list[e].C == list[k].C) ++e;
		nsh_glob::GlobArgs a = p->wargs;
		a.prob = d_list + k;
		nsh_glob::launch_ext16(a, (int)(e - k), list[k].NW, list[k].C, p->ws);
		HIPOK(hipGetLastError());
		nsh_count_launch(1);
		k = e;
	}
	HIPOK(hipEventRecord(p->wev, p->ws));
	HIPOK(hipStreamWaitEvent(p->stream, p->wev, 0));
	return 1;
fail:
	return 0;
}

// idx/n_idx: prepare only these inputs (subset runs); nullptr prepares all p->n
// capture: the launches are being recorded into a graph; the timing events are left out (graph mode
// records them around the graph launch)
static int plan_launch(nsh_ext_plan_t *p, const ExtArgs *args, const int32_t *grid, bool clear_out, const int32_t *idx,
		int32_t n_idx, const std::vector<nsh_glob::GlobProb> *wide, const nsh_glob::GlobProb *d_wide, bool capture = false)
{
	if (!capture) HIPOK(hipEventRecord(p->ev0, p->stream));
	if (n_idx > 0)
		hipLaunchKernelGGL(ext_prep_kernel, dim3(n_idx), dim3(256), 0, p->stream, p->d_prep, idx, p->d_ns, p->d_ss, p->d_as,
				p->d_rec, p->d_aas, p->d_pc);
	hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, p->stream, p->d_counter, kGroups, 0);
	if (clear_out) hipLaunchKernelGGL(clear_out_kernel, dim3((p->n + 255) / 256), dim3(256), 0, p->stream, p->d_out, p->n);
	HIPOK(hipGetLastError());
	if (!capture) HIPOK(hipEventRecord(p->ev1, p->stream));
	HIPOK(hipEventRecord(p->evf, p->stream));
	// the widest problems first: each group's persistent waves take the CUs in launch order, and a
	// group started after group 0 (the most waves) would wait for its queue to drain
	if (wide && !launch_wide(p, *wide, d_wide)) goto fail;
	for (int g = kGroups - 1; g >= 0; --g) {
		if (grid[g] <= 0) continue;
		HIPOK(hipStreamWaitEvent(p->gs[g], p->evf, 0));
		if (g == 0 && p->merged) launch_group01(args[g], grid[g], p->gs[g]);
		else launch_group(g, args[g], grid[g], p->gs[g]);
		HIPOK(hipGetLastError());
		HIPOK(hipEventRecord(p->gev[g], p->gs[g]));
		HIPOK(hipStreamWaitEvent(p->stream, p->gev[g], 0));
		nsh_count_launch(1);
	}
	if (!capture) HIPOK(hipEventRecord(p->ev2, p->stream));
	nsh_count_launch(2);
	p->ran = 1;
	return 1;
fail:
	return 0;
}

extern "C" int nsh_ext_plan_run(nsh_ext_plan_t *p)
{
	return p? plan_launch(p, p->args, p->grid, false, nullptr, p->n, &p->wide_h, p->d_wide) : 0;
}

// the same run as one hipGraph: captured from the plan's streams on the first call (and again when
// the autotuner has changed a persistent grid), replayed afterwards. The queue counters are cleared
// by a kernel inside the graph, so every replay starts from zero. The timed span of a graph run
// (nsh_ext_plan_last_ms) covers the whole graph, preparation included.
extern "C" int nsh_ext_plan_run_graph(nsh_ext_plan_t *p)
{
	if (!p) { nsh_set_error("nsh_ext_plan_run_graph: no plan"); return 0; }
	if (p->gexec && std::memcmp(p->ggrid, p->grid, sizeof(p->grid))) p->drop_graph();
	if (!p->gexec) {
		HIPOK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
		const int ok = plan_launch(p, p->args, p->grid, false, nullptr, p->n, &p->wide_h, p->d_wide, true);
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(p->stream, &g);
		if (!ok || e != hipSuccess || !g) {
			if (g) (void)hipGraphDestroy(g);
			if (ok) nsh_set_error("nsh_ext_plan_run_graph: capture failed: %s", hipGetErrorString(e));
			return 0;
		}
		p->graph = g;
		HIPOK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
		std::memcpy(p->ggrid, p->grid, sizeof(p->grid));
	}
	HIPOK(hipEventRecord(p->ev0, p->stream));
	HIPOK(hipEventRecord(p->ev1, p->stream));
	HIPOK(hipGraphLaunch(p->gexec, p->stream));
	HIPOK(hipEventRecord(p->ev2, p->stream));
	nsh_count_launch(1);
	p->ran = 1;
	return 1;
fail:
	p->drop_graph();
	return 0;
}

extern "C" int nsh_ext_plan_run_subset(nsh_ext_plan_t *p, const uint8_t *keep)
{
	if (!p || !keep) { nsh_set_error("nsh_ext_plan_run_subset: bad arguments"); return 0; }
	ExtArgs args[kGroups];
	int32_t grid[kGroups] = {};
	HIPOK(p->b_prob_sub.ensure(p->n)); HIPOK(p->b_idx_sub.ensure(p->n));
	p->d_prob_sub = p->b_prob_sub.p, p->d_idx_sub = p->b_idx_sub.p;
	HIPOK(hipStreamSynchronize(p->stream));   // the previous subset lists may still be in flight
	p->sub_h.clear(), p->idx_h.clear();
	for (int32_t k = 0; k < p->n; ++k)
		if (keep[k]) p->idx_h.push_back(k);
	for (int g = 0; g < kGroups; ++g) {
		const int32_t first = (int32_t)p->sub_h.size();
		for (int32_t k = p->grp_first[g]; k < p->grp_first[g] + p->grp_n[g]; ++k) {
			ExtProb e = p->prob_h[k];
			if (e.pair >= 0) {   // a pair runs as two single problems here
				ExtProb f = p->prob_h[e.pair];
				e.pair = -1, e.C = (e.npad + 127) / 128;
				if (keep[f.out]) p->sub_h.push_back(f);
			}
			if (keep[e.out]) p->sub_h.push_back(e);
		}
		args[g] = p->args[g];
		args[g].prob = p->d_prob_sub + first, args[g].n = (int32_t)p->sub_h.size() - first;
		grid[g] = std::min(args[g].n, p->ncu * p->wpc[g]);
	}
	p->wide_sub_h.clear();
	for (const nsh_glob::GlobProb &g : p->wide_h) if (keep[g.out]) p->wide_sub_h.push_back(g);
	if (!p->wide_sub_h.empty()) { HIPOK(p->b_wide_sub.ensure(p->wide_h.size())); p->d_wide_sub = p->b_wide_sub.p; }
	if (!p->sub_h.empty())
		HIPOK(hipMemcpyAsync(p->d_prob_sub, p->sub_h.data(), p->sub_h.size() * sizeof(ExtProb), hipMemcpyHostToDevice, p->stream));
	if (!p->idx_h.empty())
		HIPOK(hipMemcpyAsync(p->d_idx_sub, p->idx_h.data(), p->idx_h.size() * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
	if (!p->wide_sub_h.empty())
		HIPOK(hipMemcpyAsync(p->d_wide_sub, p->wide_sub_h.data(), p->wide_sub_h.size() * sizeof(nsh_glob::GlobProb),
				hipMemcpyHostToDevice, p->stream));
	return plan_launch(p, args, grid, true, p->d_idx_sub, (int32_t)p->idx_h.size(), &p->wide_sub_h, p->d_wide_sub);
fail:
	return 0;
}

extern "C" const float *nsh_ext_plan_features(nsh_ext_plan_t *p, float *host_out)
{
	if (!p) return nullptr;
	if (p->from_genome) { nsh_set_error("nsh_ext_plan_features: the plan reads its bases from a genome"); return nullptr; }
	if (p->max_al > 4096 || p->max_nl > 24000) {
		nsh_set_error("nsh_ext_plan_features: candidates limited to al <= 4096, nl <= 24000 (plan has %d, %d)", p->max_al, p->max_nl);
		return nullptr;
	}
	if (!p->d_feat) { HIPOK(p->b_feat.ensure((size_t)p->n * 64)); p->d_feat = p->b_feat.p; }
	nsh_feat::launch(p->d_prep, p->n, p->d_ns, p->d_as, p->d_pc, p->d_mat, p->d_feat, p->stream);
	HIPOK(hipGetLastError());
	if (host_out) HIPOK(hipMemcpyAsync(host_out, p->d_feat, (size_t)p->n * 64 * 4, hipMemcpyDeviceToHost, p->stream));
	HIPOK(hipStreamSynchronize(p->stream));
	nsh_count_launch(1);
	return p->d_feat;
fail:
	return nullptr;
}

extern "C" int nsh_ext_plan_sync(nsh_ext_plan_t *p)
{
	if (!p) return 0;
	HIPOK(hipStreamSynchronize(p->stream));
	return 1;
fail:
	return 0;
}

int nsh_ext_plan_query(nsh_ext_plan_t *p)
{
	return !p || hipStreamQuery(p->stream) != hipErrorNotReady;
}

extern "C" int nsh_ext_plan_results(nsh_ext_plan_t *p, nsh_ext_result_t *out)
{
	if (!p || !p->ran) { nsh_set_error("nsh_ext_plan_results: plan has not run"); return 0; }
	HIPOK(hipMemcpyAsync(p->out_h.data(), p->d_out, (size_t)16 * p->n, hipMemcpyDeviceToHost, p->stream));
	HIPOK(hipStreamSynchronize(p->stream));
	if (out)
		for (int32_t k = 0; k < p->n; ++k)
			out[k].nt_len = p->out_h[4 * k], out[k].aa_len = p->out_h[4 * k + 1], out[k].score = p->out_h[4 * k + 2];
	return 1;
fail:
	return 0;
}

extern "C" double nsh_ext_plan_cells(const nsh_ext_plan_t *p, int32_t kind)
{
	if (!p) return 0;
	if (kind == 0) return p->cells_window;
	double s = 0;
	for (int32_t k = 0; k < p->n; ++k) s += (double)p->out_h[4 * k + 3] * (kind == 2? 1 : p->al[k]);
	return s;
}

extern "C" float nsh_ext_plan_last_ms(nsh_ext_plan_t *p, int32_t which)
{
	float ms = 0;
	if (!p || !p->ran) return -1.f;
	if (hipEventElapsedTime(&ms, which == 1? p->ev1 : p->ev0, p->ev2) != hipSuccess) return -1.f;
	return ms;
}

extern "C" int nsh_ext_plan_autotune(nsh_ext_plan_t *p, int32_t reps, int32_t *waves_per_cu)
{
	if (!p || reps < 1) { nsh_set_error("nsh_ext_plan_autotune: bad arguments"); return 0; }
	int32_t *wpc = p->wpc;
	auto time_run = [&](float *best) -> int {
		*best = 1e30f;
		for (int r = 0; r < reps; ++r) {
			if (!plan_launch(p, p->args, p->grid, false, nullptr, p->n, &p->wide_h, p->d_wide)) return 0;
			if (hipEventSynchronize(p->ev2) != hipSuccess) { nsh_set_error("nsh_ext_plan_autotune: run failed"); return 0; }
			float ms = 0;
			if (hipEventElapsedTime(&ms, p->ev1, p->ev2) != hipSuccess) return 0;
			*best = std::min(*best, ms);
		}
		return 1;
	};
	{   // warm-up runs first: the clocks settle before anything is compared
		float ms;
		if (!time_run(&ms) || !time_run(&ms)) return 0;
	}
	for (int g = 0; g < kGroups; ++g) {
		const int32_t n = p->args[g].n;
		if (n <= 0) continue;
		const int32_t d = wpc[g];
		const int32_t cand[5] = { d, std::max(1, d / 2), std::max(1, 3 * d / 4), 3 * d / 2, 2 * d };
		int32_t best_w = d;
		float best_ms = 1e30f;
		for (int c = 0; c < 5; ++c) {
			if (c && (cand[c] == d || cand[c] == cand[c - 1])) continue;
			p->grid[g] = (int32_t)std::min<int64_t>(n, (int64_t)p->ncu * cand[c]);
			float ms;
			if (!time_run(&ms)) return 0;
			// the default stays unless another setting is clearly (2%) faster: run-to-run noise
			// is about 1%
			if (c == 0? true : ms < best_ms * 0.98f) best_ms = ms, best_w = cand[c];
		}
		wpc[g] = best_w;
		p->grid[g] = (int32_t)std::min<int64_t>(n, (int64_t)p->ncu * best_w);
	}
	// wavefronts per problem: two small problems per wave for every length, only up to 16,384 or
	// 65,536 rows, or never (a long problem alone on a wave runs its rows with less latency)
	if (!p->pssm_plan && p->grp_n[0] > 0) {
		const int32_t cand[4] = { INT32_MAX, 65536, 16384, 0 };
		int32_t best_r = p->pair_rows;
		float best_ms = 1e30f;
		for (int c = 0; c < 4; ++c) {
			if (!plan_requeue(p, cand[c])) return 0;
			float ms;
			if (!time_run(&ms)) return 0;
			if (c == 0? true : ms < best_ms * 0.98f) best_ms = ms, best_r = cand[c];
		}
		if (!plan_requeue(p, best_r)) return 0;
	}
	if (waves_per_cu) for (int g = 0; g < kGroups; ++g) waves_per_cu[g] = wpc[g];
	return 1;
}

extern "C" int32_t nsh_ext_plan_pair_rows(const nsh_ext_plan_t *p)
{
	return p? p->pair_rows : -1;
}

// ======================================================================================
// nsh_ext_run_noalloc: nsh_ext_plan_create + run + results in one call that allocates no device
// memory. Every device buffer of the run (the raw inputs, the row records, descriptors, the
// length-penalty table, queue counters, results) is carved out of one file-scope device array
// (`scratch`, kNoAllocScratch bytes per GPU, held until the process ends); host-side descriptors
// live in file-scope host arrays. Calls serialise on a host mutex; the device array keeps the last
// call's contents, and everything a run reads is written by that run (the queue counters are
// cleared by a fill kernel, not a memset, so the launches can also be captured in a graph).
// ======================================================================================
namespace {
constexpr size_t kNoAllocScratch = (size_t)128 << 20;   // device scratch per GPU
constexpr int32_t kNoAllocMaxProbs = 16384;               // problems per call
constexpr int32_t kNoAllocMaxNl = 262143;                 // DNA length per problem
static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

std::mutex g_na_mu;
constexpr int kMaxDev = 64;
uint8_t *g_na_ptr[kMaxDev] = {};       // device address of `scratch` on each GPU
ExtProb g_na_prob[kNoAllocMaxProbs];
PrepProb g_na_prep[kNoAllocMaxProbs];
int32_t g_na_order[kNoAllocMaxProbs];
int32_t g_na_out[4 * kNoAllocMaxProbs];
int32_t g_na_pen[kNoAllocMaxNl + 2];
PrepConst g_na_pc;
size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
} // namespace

extern "C" int nsh_ext_run_noalloc(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl,
		const char *as, const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt,
		nsh_ext_result_t *out, void *stream_)
{
	hipStream_t stream = (hipStream_t)stream_;
	if (n < 1 || !ns || !ns_off || !nl || !as || !as_off || !al || !opt || !out) {
		nsh_set_error("nsh_ext_run_noalloc: bad arguments"); return 0;
	}
	if (n > kNoAllocMaxProbs) { nsh_set_error("nsh_ext_run_noalloc: %d problems; the limit is %d", n, kNoAllocMaxProbs); return 0; }
	if (!nsh_ensure_device()) return 0;
	if (opt->asize != 22) { nsh_set_error("asize %d not supported (22 expected)", opt->asize); return 0; }
	if (opt->pssm) { nsh_set_error("nsh_ext_run_noalloc: PSSM rows are not supported (matrix scores only)"); return 0; }
	if (opt->ge < 0 || opt->go < 0 || (int64_t)opt->ge * CMAX * 64 > 32767) {
		nsh_set_error("gap penalties go=%d ge=%d out of the supported range", opt->go, opt->ge); return 0;
	}
	for (int t = 0; t < 64; ++t)
		if (opt->codon[t] >= 22) { nsh_set_error("codon table entry %d out of range", t); return 0; }
	for (int t = 0; t < 256; ++t)
		if (opt->aa20[t] >= 22 || opt->nt4[t] > 4) { nsh_set_error("alphabet table entry %d out of range", t); return 0; }
	std::lock_guard<std::mutex> lk(g_na_mu);
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) { nsh_set_error("nsh_ext_run_noalloc: no current device"); return 0; }
	if (!g_na_ptr[dev]) {   // once per device, with that device current
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr) { nsh_set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)); return 0; }
		g_na_ptr[dev] = (uint8_t*)ptr;
	}
	uint8_t *base = g_na_ptr[dev];
	const bool global = !dirs && !(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
	int64_t span_ns = 0, span_as = 0, l_rec = 0, l_aa = 0;
	int32_t max_nl = 2, n_grp[kGroups] = {};
	for (int32_t k = 0; k < n; ++k) {
		if (al[k] < 1 || nl[k] < 0 || ns_off[k] < 0 || as_off[k] < 0) {
			nsh_set_error("nsh_ext_run_noalloc: problem %d: nl=%d al=%d", k, nl[k], al[k]); return 0;
		}
		if (nl[k] > kNoAllocMaxNl) { nsh_set_error("nsh_ext_run_noalloc: problem %d: nl=%d exceeds %d", k, nl[k], kNoAllocMaxNl); return 0; }
		const int32_t flag = dirs? (dirs[k]? NS_F_EXT_LEFT : NS_F_EXT_RIGHT) : (opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT));
		const int32_t w = nsh_host_stripe_width(global? 0 : flag, al[k]);
		const int32_t npad = (al[k] + w - 1) / w * w;
		const int32_t C = std::max(1, (npad + 127) / 128);
		if (C > CMAX) { nsh_set_error("nsh_ext_run_noalloc: problem %d: padded protein length %d exceeds %d", k, npad, 128 * CMAX); return 0; }
		ExtProb &e = g_na_prob[k];
		e.nl = nl[k], e.al = al[k], e.npad = npad, e.io = opt->io, e.out = k, e.C = C;
		e.left = (flag & NS_F_EXT_LEFT) != 0, e.mode = global;
		e.rec_off = l_rec, e.aa_off = l_aa, e.pssm_off = -1;
		PrepProb &r = g_na_prep[k];
		r.ns_off = ns_off[k], r.as_off = as_off[k], r.rec_off = l_rec, r.aa_off = l_aa;
		r.nl = nl[k], r.al = al[k], r.left = e.left, r.io = opt->io, r.has_ss = 0;
		l_rec += nl[k] + 1, l_aa += al[k] + 1;
		span_ns = std::max(span_ns, ns_off[k] + nl[k]), span_as = std::max(span_as, as_off[k] + al[k]);
		max_nl = std::max(max_nl, nl[k]);
		++n_grp[group_of(C)];
	}
	// scratch layout
	size_t o = 0;
	const size_t o_ns = o; o = na_align(o + span_ns + 1);
	const size_t o_as = o; o = na_align(o + span_as + 1);
	const size_t o_aas = o; o = na_align(o + l_aa);
	const size_t o_rec = o; o = na_align(o + (size_t)l_rec * 4);
	const size_t o_mat = o; o = na_align(o + 22 * 22);
	const size_t o_pen = o; o = na_align(o + (size_t)(max_nl + 2) * 4);
	const size_t o_cnt = o; o = na_align(o + kGroups * 4);
	const size_t o_out = o; o = na_align(o + (size_t)n * 16);
	const size_t o_prob = o; o = na_align(o + (size_t)n * sizeof(ExtProb));
	const size_t o_prep = o; o = na_align(o + (size_t)n * sizeof(PrepProb));
	const size_t o_pc = o; o = na_align(o + sizeof(PrepConst));
	if (o > kNoAllocScratch) {
		nsh_set_error("nsh_ext_run_noalloc: the inputs need %zu bytes of device scratch; the limit is %zu", o, kNoAllocScratch);
		return 0;
	}
	for (int32_t d = 0; d < max_nl + 2; ++d) g_na_pen[d] = nsh_ie_penalty(opt->ie_coef, d);
	std::memcpy(g_na_pc.nt4, opt->nt4, 256), std::memcpy(g_na_pc.aa20, opt->aa20, 256), std::memcpy(g_na_pc.codon, opt->codon, 64);
	for (int t = 0; t < 6; ++t) g_na_pc.sp[t] = opt->sp[t];
	g_na_pc.sp_null = opt->sp_null_bonus;
	// queue order as in a plan: by column-count group, longest work first (index breaks ties)
	for (int32_t k = 0; k < n; ++k) g_na_order[k] = k;
	std::sort(g_na_order, g_na_order + n, [](int32_t x, int32_t y) {
		const ExtProb &a = g_na_prob[x], &b = g_na_prob[y];
		const int ga = group_of(a.C), gb = group_of(b.C);
		if (ga != gb) return ga < gb;
		const int64_t ca = (int64_t)a.nl * a.C, cb = (int64_t)b.nl * b.C;
		return ca != cb? ca > cb : x < y; });
	ExtProb *d_prob = (ExtProb*)(base + o_prob);
	int ok = 1;
	auto cp = [&](void *dst, const void *src, size_t bytes) {
		if (ok && bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) ok = 0;
	};
	cp(base + o_ns, ns, (size_t)span_ns);
	cp(base + o_as, as, (size_t)span_as);
	cp(base + o_mat, opt->sc, 22 * 22);
	cp(base + o_pen, g_na_pen, (size_t)(max_nl + 2) * 4);
	cp(base + o_prep, g_na_prep, (size_t)n * sizeof(PrepProb));
	cp(base + o_pc, &g_na_pc, sizeof(PrepConst));
	for (int32_t k = 0; k < n; ++k) cp(d_prob + k, &g_na_prob[g_na_order[k]], sizeof(ExtProb));
	if (!ok) { nsh_set_error("nsh_ext_run_noalloc: copy to the device failed"); return 0; }
	nsh_prep::launch((const PrepProb*)(base + o_prep), n, base + o_ns, nullptr, base + o_as, (uint32_t*)(base + o_rec),
			base + o_aas, (const PrepConst*)(base + o_pc), stream);
	int32_t *d_cnt = (int32_t*)(base + o_cnt);
	hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, d_cnt, kGroups, 0);
	int ncu = 0;
	(void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
	if (ncu <= 0) ncu = 256;
	int64_t first = 0;
	for (int g = 0; g < kGroups; ++g) {
		ExtArgs a = {};
		a.prob = d_prob + first, a.n = n_grp[g], a.pbase = d_prob;
		first += n_grp[g];
		if (a.n == 0) continue;
		a.rec = (const uint32_t*)(base + o_rec), a.aas = base + o_aas, a.pssm = nullptr, a.mat = (const int8_t*)(base + o_mat);
		a.asize = 22, a.pen = (const int32_t*)(base + o_pen), a.pen_n = max_nl + 2;
		a.go = opt->go, a.ge = opt->ge, a.fs = opt->fs, a.xdrop = opt->xdrop, a.eb = opt->end_bonus;
		a.counter = d_cnt + g, a.out = (int32_t*)(base + o_out);
		const int grid = (int)std::min<int64_t>(a.n, (int64_t)ncu * kGroupWavesPerCU[g]);
		launch_group(g, a, grid, stream);   // one stream: the groups run one after the other
	}
	hipError_t e = hipGetLastError();
	if (e == hipSuccess) e = hipMemcpyAsync(g_na_out, base + o_out, (size_t)n * 16, hipMemcpyDeviceToHost, stream);
	if (e == hipSuccess) e = hipStreamSynchronize(stream);
	if (e != hipSuccess) { nsh_set_error("nsh_ext_run_noalloc: %s", hipGetErrorString(e)); return 0; }
	nsh_count_launch(2 + kGroups);
	for (int32_t k = 0; k < n; ++k)
		out[k].nt_len = g_na_out[4 * k], out[k].aa_len = g_na_out[4 * k + 1], out[k].score = g_na_out[4 * k + 2];
	return 1;
}

extern "C" void nsh_ext_plan_destroy(nsh_ext_plan_t *p)
{
	if (!p) return;
	shell_release(p);   // waits for the plan's work; its resources go back to the pool (or are freed)
}
