// C-ABI entry points of libnasw_hip: device bring-up, error reporting, the ns_* DP calls and
// the reference's batched-extension API (nasw.h:98-212). All DP work is dispatched to the
// GPU engines (ext_dp.hip, global_dp.hip); nothing here computes a DP on the host.
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "nsh_internal.h"

// ---------------------------------------------------------------------------------------
// kalloc fall-backs (kalloc.h:14-32). Weak: a caller that links the reference's kalloc.c
// (per-thread arenas) overrides them; here km must be NULL (plain libc heap).
// ---------------------------------------------------------------------------------------
extern "C" {
__attribute__((weak)) void *kmalloc(void *km, size_t size)
{
	if (km) { std::fprintf(stderr, "[nasw_hip] kmalloc: a non-NULL km needs the caller's kalloc\n"); std::abort(); }
	return std::malloc(size);
}
__attribute__((weak)) void *krealloc(void *km, void *ptr, size_t size)
{
	if (km) { std::fprintf(stderr, "[nasw_hip] krealloc: a non-NULL km needs the caller's kalloc\n"); std::abort(); }
	return std::realloc(ptr, size);
}
__attribute__((weak)) void *kcalloc(void *km, size_t count, size_t size)
{
	if (km) { std::fprintf(stderr, "[nasw_hip] kcalloc: a non-NULL km needs the caller's kalloc\n"); std::abort(); }
	return std::calloc(count, size);
}
__attribute__((weak)) void kfree(void *km, void *ptr)
{
	if (km) { std::fprintf(stderr, "[nasw_hip] kfree: a non-NULL km needs the caller's kalloc\n"); std::abort(); }
	std::free(ptr);
}
}

// ---------------------------------------------------------------------------------------
// device state, errors, counters
// ---------------------------------------------------------------------------------------
namespace {
std::mutex g_err_mu, g_init_mu;
char g_err[512] = "no error";
char g_name[256] = "";
int g_dev_state = 0;        // 0 unknown, 1 ready, -1 unusable
int g_device = 0;
std::atomic<uint64_t> g_dispatch{0}, g_launch{0}, g_ext_launch{0}, g_ext_req{0};
std::atomic<uint64_t> g_h2d_ns{0}, g_kernel_ns{0}, g_d2h_ns{0};
} // namespace

void nsh_set_error(const char *fmt, ...)
{
	std::lock_guard<std::mutex> lk(g_err_mu);
	va_list ap;
	va_start(ap, fmt);
	std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

static int nsh_open_device(int device)
{
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n <= 0) { nsh_set_error("no HIP device: %s", hipGetErrorString(e)); return 0; }
	if (device < 0 || device >= n) { nsh_set_error("HIP device %d out of range (%d devices)", device, n); return 0; }
	hipDeviceProp_t prop;
	if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) { nsh_set_error("hipGetDeviceProperties: %s", hipGetErrorString(e)); return 0; }
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
		nsh_set_error("device %d is %s; this library carries gfx950 code only", device, prop.gcnArchName);
		return 0;
	}
	if ((e = hipSetDevice(device)) != hipSuccess) { nsh_set_error("hipSetDevice: %s", hipGetErrorString(e)); return 0; }
	std::snprintf(g_name, sizeof(g_name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
	g_device = device;
	return 1;
}

int nsh_ensure_device()
{
	std::lock_guard<std::mutex> lk(g_init_mu);
	if (g_dev_state == 0) {
		const char *s = std::getenv("NS_GPU_TEST_DEVICE");
		g_dev_state = nsh_open_device(s? std::atoi(s) : 0)? 1 : -1;
	} else if (g_dev_state == 1) {
		(void)hipSetDevice(g_device); // make the library's device current on this thread
	}
	return g_dev_state == 1;
}

void nsh_count_launch(uint64_t n) { g_launch += n; }
void nsh_add_times(double h2d, double kernel, double d2h)
{
	g_h2d_ns += (uint64_t)(h2d * 1e9), g_kernel_ns += (uint64_t)(kernel * 1e9), g_d2h_ns += (uint64_t)(d2h * 1e9);
}

// the reference CPU kernel's SIMD width for a request (nasw-sse.c:744-750): 16 for extension with
// al >= 16 on AVX2 hosts, else 8. Extension results depend on it (padded columns feed the row max).
int32_t nsh_host_stripe_width(int32_t flag, int32_t al)
{
	static int avx2 = -1;
	if (avx2 < 0) {
		const char *s = std::getenv("NSH_STRIPE_WIDTH");
		if (s) avx2 = std::atoi(s) == 16;
		else {
#if defined(__x86_64__) || defined(__i386__)
			avx2 = __builtin_cpu_supports("avx2")? 1 : 0;
#else
			avx2 = 0;
#endif
		}
	}
	return (flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) && al >= 16 && avx2? 16 : 8;
}

// nasw-sse.c:341-349: log2 as exponent + quadratic in the mantissa (valid for x >= 2), then
// (int)(ie_coef * log2 + .5f). Each float operation is its own statement so that no fused
// multiply-add can change the rounding (this file is also compiled with -ffp-contract=off).
int32_t nsh_ie_penalty(float ie_coef, int32_t d)
{
	if (d < 2) return 0;
	union { float f; uint32_t i; } z;
	z.f = (float)d;
	float lg = (float)(int32_t)((z.i >> 23) & 255) - 128;
	z.i &= ~(255u << 23);
	z.i += 127u << 23;
	const float m = z.f;
	float poly = -0.34484843f * m;
	poly = poly + 2.02466578f;
	poly = poly * m;
	poly = poly - 0.67487759f;
	lg = lg + poly;
	float t = ie_coef * lg;
	t = t + .5f;
	return (int32_t)t;
}

// ---------------------------------------------------------------------------------------
// GPU global-DP API (nasw.h:140-152)
// ---------------------------------------------------------------------------------------
extern "C" int ns_gpu_init(int32_t device)
{
	std::lock_guard<std::mutex> lk(g_init_mu);
	if (g_dev_state == 1 && device == g_device) return 1;
	g_dev_state = nsh_open_device(device)? 1 : -1;
	return g_dev_state == 1;
}
extern "C" const char *ns_gpu_error(void) { return g_err; }
extern "C" const char *ns_gpu_name(void) { return g_name; }
extern "C" uint64_t ns_gpu_dispatch_count(void) { return g_dispatch.load(); }
extern "C" uint64_t ns_gpu_launch_count(void) { return g_launch.load(); }
extern "C" double ns_gpu_h2d_seconds(void) { return g_h2d_ns.load() * 1e-9; }
extern "C" double ns_gpu_kernel_seconds(void) { return g_kernel_ns.load() * 1e-9; }
extern "C" double ns_gpu_d2h_seconds(void) { return g_d2h_ns.load() * 1e-9; }
extern "C" uint64_t ns_ext_gpu_launch_count(void) { return g_ext_launch.load(); }
extern "C" uint64_t ns_ext_gpu_request_count(void) { return g_ext_req.load(); }

extern "C" int32_t ns_ext_stripe_width(const ns_opt_t *opt, int32_t al) { return nsh_host_stripe_width(opt->flag, al); }
extern "C" int32_t ns_ext_ie_penalty(const ns_opt_t *opt, int32_t d) { return nsh_ie_penalty(opt->ie_coef, d); }

// one extension request through a single-request plan
static int nsh_ext_one(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	nsh_ext_req q;
	q.ns = ns, q.nl = nl, q.as = as, q.al = al, q.ss = ss;
	q.left = (opt->flag & NS_F_EXT_LEFT) != 0, q.io = opt->io;
	q.mode = (opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) == 0;
	q.pssm = opt->pssm;
	const int32_t w = nsh_host_stripe_width(opt->flag, al);
	q.npad = (al + w - 1) / w * w;
	nsh_ext_plan_t *p = nsh_ext_plan_build(1, &q, opt);
	if (!p) return 0;
	nsh_ext_result_t res;
	int ok = nsh_ext_plan_run(p) && nsh_ext_plan_results(p, &res);
	nsh_ext_plan_destroy(p);
	if (!ok) return 0;
	r->n_cigar = 0, r->nt_len = res.nt_len, r->aa_len = res.aa_len, r->score = res.score;
	++g_dispatch, ++g_ext_launch;
	return 1;
}

static void nsh_fail_result(ns_rst_t *r, int32_t nl, int32_t al)
{
	r->n_cigar = 0, r->nt_len = nl, r->aa_len = al, r->score = INT32_MIN;
}

static void nsh_global_dispatch(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, int32_t is32, ns_rst_t *r)
{
	int ok;
	// int16 score-only (extension or plain global) runs on the extension engine (nasw-sse.c:360-459)
	const bool ext = (opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) != 0;
	if (!is32 && (ext || !(opt->flag & NS_F_CIGAR))) ok = nsh_ext_one(ns, nl, as, al, opt, ss, r);
	else ok = nsh_global_run(km, ns, nl, as, al, opt, ss, is32, r);
	if (!ok) {
		std::fprintf(stderr, "[nasw_hip] GPU DP failed (nl=%d al=%d): %s\n", nl, al, ns_gpu_error());
		nsh_fail_result(r, nl, al);
	}
}

extern "C" void ns_global_gs16b(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	nsh_global_dispatch(km, ns, nl, as, al, opt, ss, 0, r);
}
extern "C" void ns_global_gs32b(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	nsh_global_dispatch(km, ns, nl, as, al, opt, ss, 1, r);
}
extern "C" void ns_global_gs16(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, ns_rst_t *r)
{
	ns_global_gs16b(km, ns, nl, as, al, opt, nullptr, r);
}
extern "C" void ns_global_gs32(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, ns_rst_t *r)
{
	ns_global_gs32b(km, ns, nl, as, al, opt, nullptr, r);
}

// nasw-sse.c:980-999: int32 when the best possible score could exceed the int16 range
extern "C" void ns_global_gs_auto(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	int32_t max_sub = 0;
	if (opt->pssm) { for (int32_t i = 0; i < al * opt->asize; ++i) if (max_sub < opt->pssm[i]) max_sub = opt->pssm[i]; }
	else { for (int32_t i = 0; i < opt->asize * opt->asize; ++i) if (max_sub < opt->sc[i]) max_sub = opt->sc[i]; }
	const int64_t max_score = (int64_t)al * max_sub + opt->end_bonus;
	if (!(opt->flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) && max_score > 30000) ns_global_gs32b(km, ns, nl, as, al, opt, ss, r);
	else ns_global_gs16b(km, ns, nl, as, al, opt, ss, r);
}

// ns_gpu_submit/wait/global (nasw.h:140-151, nasw-cuda.cu:1068-1145): the int16 global DP with
// CIGAR. Like the reference, a request is refused (NULL / 0) when nl < 3, al < 1, asize < 1 or
// the best possible score al * max(sc) + end_bonus exceeds 30000 (the caller then takes the int32
// path). The extension flags are ignored and the CIGAR is always produced; the splice-score track
// is honoured and opt->pssm is ignored (the reference's kernel scores with opt->sc only,
// nasw-cuda.cu:1068-1110), so the result equals ns_global_gs16b with NS_F_CIGAR and no PSSM.
// Submission copies the inputs and queues the problem in the global DP engine (global_dp.hip),
// which batches the submissions of all threads; each submission is consumed exactly once by
// ns_gpu_wait, which copies the CIGAR into km on the calling thread.
struct ns_gpu_request_s { nsh_glob_req *g; void *km; };

static int nsh_gpu_global_accepts(int32_t nl, int32_t al, const ns_opt_t *opt)
{
	if (nl < 3 || al < 1 || opt->asize < 1) return 0;
	int32_t max_sub = 0;
	for (int32_t i = 0; i < opt->asize * opt->asize; ++i) if (max_sub < opt->sc[i]) max_sub = opt->sc[i];
	return (int64_t)al * max_sub + opt->end_bonus <= 30000;
}

// the request ns_gpu_submit queues: int16, CIGAR, no extension, matrix scores
static nsh_glob_req *nsh_gpu_global_submit(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss)
{
	if (!nsh_ensure_device() || !nsh_gpu_global_accepts(nl, al, opt)) return nullptr;
	ns_opt_t o = *opt;
	o.flag = (o.flag | NS_F_CIGAR) & ~(NS_F_GPU | NS_F_EXT_LEFT | NS_F_EXT_RIGHT);
	o.pssm = nullptr;
	return nsh_glob_submit(ns, nl, as, al, &o, ss, 0);
}

extern "C" ns_gpu_request_t *ns_gpu_submit(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss)
{
	nsh_glob_req *g = nsh_gpu_global_submit(ns, nl, as, al, opt, ss);
	if (!g) return nullptr;
	ns_gpu_request_t *q = (ns_gpu_request_t*)std::malloc(sizeof(ns_gpu_request_t));
	if (!q) { ns_rst_t tmp; ns_rst_init(&tmp); (void)nsh_glob_wait(g, nullptr, &tmp); std::free(tmp.cigar); return nullptr; }
	q->g = g, q->km = km;
	return q;
}

extern "C" int ns_gpu_wait(ns_gpu_request_t *q, ns_rst_t *r)
{
	if (!q) return 0;
	const int ok = nsh_glob_wait(q->g, q->km, r);   // appends to the CIGAR r may already hold
	std::free(q);
	if (ok) ++g_dispatch;
	return ok;
}

extern "C" int ns_gpu_global(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	nsh_glob_req *g = nsh_gpu_global_submit(ns, nl, as, al, opt, ss);
	if (!g) return 0;
	r->n_cigar = 0;
	const int ok = nsh_glob_wait(g, km, r);
	if (ok) ++g_dispatch;
	return ok;
}

// ---------------------------------------------------------------------------------------
// batched extension API (nasw.h:196-212). A flush turns the requests added since the last
// flush into one generation: requests with identical options (io, direction and the PSSM rows
// may differ) share one device plan; the plans run asynchronously on their own streams.
// ---------------------------------------------------------------------------------------
namespace {
struct ExtGen {
	int32_t base = 0, n = 0, n_got = 0;
	std::vector<ns_rst_t> rst;
	std::vector<nsh_ext_plan_t*> plans;
	std::vector<std::vector<int32_t>> members; // request index (within the generation) per plan slot
	bool finalized = false, failed = false;
};
struct ExtReq { const char *ns, *as; const uint8_t *ss; int32_t nl, al; ns_opt_t opt; };

bool opt_same(const ns_opt_t &a, const ns_opt_t &b)
{
	return a.go == b.go && a.ge == b.ge && a.fs == b.fs && a.xdrop == b.xdrop && a.end_bonus == b.end_bonus &&
		a.asize == b.asize && a.ie_coef == b.ie_coef && a.sp_null_bonus == b.sp_null_bonus &&
		std::memcmp(a.sp, b.sp, sizeof(a.sp)) == 0 &&
		(a.sc == b.sc || std::memcmp(a.sc, b.sc, (size_t)a.asize * a.asize) == 0) &&
		(a.nt4 == b.nt4 || std::memcmp(a.nt4, b.nt4, 256) == 0) &&
		(a.aa20 == b.aa20 || std::memcmp(a.aa20, b.aa20, 256) == 0) &&
		(a.codon == b.codon || std::memcmp(a.codon, b.codon, 64) == 0);
}
} // namespace

struct ns_ext_batch_s {
	void *km = nullptr;
	int32_t next_id = 0;
	std::vector<ExtReq> pending;
	std::vector<ExtGen*> gens;
};

extern "C" ns_ext_batch_t *ns_ext_batch_init(void *km)
{
	if (!nsh_ensure_device()) return nullptr;
	ns_ext_batch_t *b = new ns_ext_batch_s();
	b->km = km;
	return b;
}

extern "C" int32_t ns_ext_batch_add(ns_ext_batch_t *b, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss)
{
	ExtReq q;
	q.ns = ns, q.as = as, q.ss = ss, q.nl = nl, q.al = al, q.opt = *opt;
	q.opt.flag &= ~(NS_F_CIGAR | NS_F_GPU);
	b->pending.push_back(q);
	return b->next_id++;
}

static void gen_finalize(ExtGen *g)
{
	if (g->finalized) return;
	std::vector<nsh_ext_result_t> res;
	for (size_t s = 0; s < g->plans.size(); ++s) {
		const std::vector<int32_t> &mem = g->members[s];
		res.resize(mem.size());
		nsh_ext_plan_t *p = g->plans[s];
		const int ok = p && nsh_ext_plan_results(p, res.data());
		for (size_t k = 0; k < mem.size(); ++k) {
			ns_rst_t &r = g->rst[mem[k]];
			ns_rst_init(&r);
			if (ok) r.nt_len = res[k].nt_len, r.aa_len = res[k].aa_len, r.score = res[k].score;
			else r.score = INT32_MIN, g->failed = true;
		}
		nsh_ext_plan_destroy(p);
		g->plans[s] = nullptr;
	}
	if (g->failed) std::fprintf(stderr, "[nasw_hip] extension batch failed on the GPU: %s\n", ns_gpu_error());
	g->finalized = true;
}

static void batch_reap(ns_ext_batch_t *b)
{
	std::vector<ExtGen*> keep;
	for (ExtGen *g : b->gens) {
		if (g->finalized && g->n_got >= g->n) delete g;
		else keep.push_back(g);
	}
	b->gens.swap(keep);
}

extern "C" int32_t ns_ext_batch_flush(ns_ext_batch_t *b)
{
	if (!b) return 0;
	batch_reap(b);
	const int32_t n = (int32_t)b->pending.size();
	if (n == 0) return 0;
	ExtGen *g = new ExtGen();
	g->base = b->next_id - n, g->n = n;
	g->rst.resize(n);
	// group by options (io and direction are per request)
	std::vector<int32_t> group(n, -1);
	std::vector<const ns_opt_t*> reps;
	for (int32_t k = 0; k < n; ++k) {
		int32_t s = 0;
		for (; s < (int32_t)reps.size(); ++s) if (opt_same(*reps[s], b->pending[k].opt)) break;
		if (s == (int32_t)reps.size()) reps.push_back(&b->pending[k].opt);
		group[k] = s;
	}
	g->members.resize(reps.size());
	for (int32_t k = 0; k < n; ++k) g->members[group[k]].push_back(k);
	for (size_t s = 0; s < reps.size(); ++s) {
		const std::vector<int32_t> &mem = g->members[s];
		std::vector<nsh_ext_req> req(mem.size());
		for (size_t k = 0; k < mem.size(); ++k) {
			const ExtReq &x = b->pending[mem[k]];
			nsh_ext_req &q = req[k];
			q.ns = x.ns, q.nl = x.nl, q.as = x.as, q.al = x.al, q.ss = x.ss;
			q.left = (x.opt.flag & NS_F_EXT_LEFT) != 0, q.io = x.opt.io;
			q.pssm = x.opt.pssm;   // per request: requests that differ only in their PSSM share a plan
			const int32_t w = nsh_host_stripe_width(x.opt.flag, x.al);
			q.npad = (x.al + w - 1) / w * w;
		}
		nsh_ext_plan_t *p = nsh_ext_plan_build((int32_t)req.size(), req.data(), reps[s]);
		if (p && !nsh_ext_plan_run(p)) nsh_ext_plan_destroy(p), p = nullptr;
		g->plans.push_back(p);
		if (p) g_ext_launch += 1;
	}
	g_ext_req += n;
	g_dispatch += n;
	b->pending.clear();
	b->gens.push_back(g);
	return n;
}

static ExtGen *batch_find(ns_ext_batch_t *b, int32_t id)
{
	for (ExtGen *g : b->gens)
		if (id >= g->base && id < g->base + g->n) return g;
	return nullptr;
}

static bool gen_done(ExtGen *g)
{
	if (g->finalized) return true;
	for (nsh_ext_plan_t *p : g->plans)
		if (p && nsh_ext_plan_query(p) == 0) return false;
	return true;
}

extern "C" int ns_ext_batch_ready(ns_ext_batch_t *b, int32_t id)
{
	ExtGen *g = batch_find(b, id);
	if (!g) return 0;
	if (!gen_done(g)) return 0;
	gen_finalize(g);
	return 1;
}

extern "C" void ns_ext_batch_wait(ns_ext_batch_t *b)
{
	if (!b) return;
	// block until some unfinalized generation completes (or none is outstanding)
	for (ExtGen *g : b->gens) {
		if (g->finalized) continue;
		for (nsh_ext_plan_t *p : g->plans) if (p) nsh_ext_plan_sync(p);
		return;
	}
}

extern "C" const ns_rst_t *ns_ext_batch_get(ns_ext_batch_t *b, int32_t id)
{
	ExtGen *g = batch_find(b, id);
	if (!g) return nullptr;
	if (!g->finalized) {
		for (nsh_ext_plan_t *p : g->plans) if (p) nsh_ext_plan_sync(p);
		gen_finalize(g);
	}
	++g->n_got;
	return &g->rst[id - g->base];
}

extern "C" void ns_ext_batch_destroy(ns_ext_batch_t *b)
{
	if (!b) return;
	for (ExtGen *g : b->gens) {
		for (nsh_ext_plan_t *p : g->plans) if (p) nsh_ext_plan_destroy(p);
		delete g;
	}
	delete b;
}

// ---------------------------------------------------------------------------------------
// ns_ext_prep (nasw-sse.c:752-756): host-side preparation exposed for callers that build
// their own device inputs; one km block holding nas[nl], aas[al], donor[nl+1], acceptor[nl+1].
// ---------------------------------------------------------------------------------------
extern "C" uint8_t *ns_ext_prep(void *km, const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt,
		const uint8_t *ss, uint8_t **aas_, int8_t **donor_, int8_t **acc_)
{
	uint8_t *nas = (uint8_t*)kmalloc(km, (size_t)nl + al + (size_t)(nl + 1) * 2);
	uint8_t *aas = nas + nl;
	int8_t *donor = (int8_t*)aas + al, *acc = donor + nl + 1;
	*aas_ = aas, *donor_ = donor, *acc_ = acc;
	nsh_prep_host(ns, nl, as, al, opt, ss, (opt->flag & NS_F_EXT_LEFT) != 0, nas, aas, donor, acc);
	return nas;
}

extern "C" int nsh_device_synchronize(void)
{
	if (!nsh_ensure_device()) return 0;
	return hipDeviceSynchronize() == hipSuccess;
}

extern "C" const char *nsh_build_info(void)
{
	return "libnasw_hip: gfx950; ext DP packed-int16 row-synchronous wave64 kernels";
}
