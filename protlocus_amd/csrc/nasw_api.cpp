// C-ABI entry points of libnasw_hip: device bring-up, error reporting, the ns_* DP calls and
// the reference's batched-extension API (nasw.h:98-212). All DP work is dispatched to the
// GPU engines (ext_dp.hip, global_dp.hip); nothing here computes a DP on the host.
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <thread>
#include <deque>
#include <condition_variable>
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

// one extension request, queued in the shared extension executor (below) and waited for
static int nsh_ext_one(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r);

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
// batched extension API (nasw.h:196-212) and single extension calls, on one shared executor.
//
// A flush turns the requests added since the last flush into a generation and queues it; a
// synchronous ns_global_gs16b(EXT) call queues a generation of one. One executor thread takes every
// queued generation of every batch and thread at once (after a 1-ms window, at once when a caller
// is waiting, or at 65,536 queued requests), groups their requests by options (io, direction and
// the PSSM rows may differ within a group) and runs one device plan per group; plans come from a
// pool, so a dispatch allocates nothing once the pool has grown to the workload. Inputs stay the
// caller's until their results are fetched (nasw.h:196-209), as the reference requires.
// ---------------------------------------------------------------------------------------
namespace {
struct ExtReq { const char *ns, *as; const uint8_t *ss; int32_t nl, al; ns_opt_t opt; nsh_ext_result_t res; int ok; };
struct ExtGen {
	std::vector<ExtReq> req;
	std::vector<ns_rst_t> rst;    // filled when done
	int32_t base = 0, n_got = 0;
	int done = 0;
};

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

class ExtExec {
public:
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<ExtGen*> queue;
	size_t queued_req = 0;
	int n_waiting = 0;
	bool started = false;

	void submit(ExtGen *g)
	{
		std::lock_guard<std::mutex> lk(mu);
		if (!started) { started = true; std::thread([this] { loop(); }).detach(); }
		queue.push_back(g);
		queued_req += g->req.size();
		cv_work.notify_one();
	}
	bool is_done(ExtGen *g)
	{
		std::lock_guard<std::mutex> lk(mu);
		return g->done != 0;
	}
	// blocks until pred() holds (checked under the lock whenever a generation completes)
	template <class P> void wait_until(P pred)
	{
		std::unique_lock<std::mutex> lk(mu);
		if (pred()) return;
		++n_waiting;
		cv_work.notify_one();   // a waiter ends the batching window
		cv_done.wait(lk, pred);
		--n_waiting;
	}
	void loop()
	{
		for (;;) {
			std::vector<ExtGen*> gens;
			{
				std::unique_lock<std::mutex> lk(mu);
				cv_work.wait(lk, [&] { return !queue.empty(); });
				const auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(1);
				while (n_waiting == 0 && queued_req < 65536 && cv_work.wait_until(lk, until) != std::cv_status::timeout) {}
				gens.assign(queue.begin(), queue.end());
				queue.clear();
				queued_req = 0;
			}
			run(gens);
			{
				std::lock_guard<std::mutex> lk(mu);
				for (ExtGen *g : gens) g->done = 1;
			}
			cv_done.notify_all();
		}
	}
	void run(std::vector<ExtGen*> &gens)
	{
		std::vector<ExtReq*> all;
		for (ExtGen *g : gens) for (ExtReq &r : g->req) all.push_back(&r);
		std::vector<char> taken(all.size(), 0);
		bool failed = false;
		for (size_t a = 0; a < all.size(); ++a) {
			if (taken[a]) continue;
			std::vector<ExtReq*> grp;
			std::vector<nsh_ext_req> req;
			for (size_t b = a; b < all.size(); ++b) {
				if (taken[b] || !opt_same(all[a]->opt, all[b]->opt)) continue;
				taken[b] = 1;
				const ExtReq &x = *all[b];
				nsh_ext_req q;
				q.ns = x.ns, q.nl = x.nl, q.as = x.as, q.al = x.al, q.ss = x.ss;
				q.left = (x.opt.flag & NS_F_EXT_LEFT) != 0, q.io = x.opt.io;
				q.mode = (x.opt.flag & (NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) == 0;
				q.pssm = x.opt.pssm;
				const int32_t w = nsh_host_stripe_width(x.opt.flag, x.al);
				q.npad = (x.al + w - 1) / w * w;
				// a request beyond the device limits fails alone; the rest of its group still runs
				if (!nsh_ext_req_fits(q, (int32_t)b)) {
					all[b]->ok = 0;
					std::fprintf(stderr, "[nasw_hip] extension request not run (nl=%d al=%d): %s\n", x.nl, x.al, ns_gpu_error());
					continue;
				}
				grp.push_back(all[b]), req.push_back(q);
			}
			if (grp.empty()) continue;
			std::vector<nsh_ext_result_t> res(grp.size());
			nsh_ext_plan_t *p = nsh_ext_plan_build((int32_t)req.size(), req.data(), &grp[0]->opt);
			const int ok = p && nsh_ext_plan_run(p) && nsh_ext_plan_results(p, res.data());
			if (p) nsh_ext_plan_destroy(p), g_ext_launch += 1;
			for (size_t k = 0; k < grp.size(); ++k) grp[k]->ok = ok, grp[k]->res = res[k];
			failed |= !ok;
		}
		if (failed) std::fprintf(stderr, "[nasw_hip] extension batch failed on the GPU: %s\n", ns_gpu_error());
		for (ExtGen *g : gens) {
			g->rst.resize(g->req.size());
			for (size_t k = 0; k < g->req.size(); ++k) {
				ns_rst_t &r = g->rst[k];
				ns_rst_init(&r);
				const ExtReq &x = g->req[k];
				if (x.ok) r.nt_len = x.res.nt_len, r.aa_len = x.res.aa_len, r.score = x.res.score;
				else r.nt_len = x.nl, r.aa_len = x.al, r.score = INT32_MIN;
			}
		}
	}
};

ExtExec &ext_exec()
{
	static ExtExec *e = new ExtExec();   // never destroyed: its thread lives until the process ends
	return *e;
}

ExtReq make_req(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss)
{
	ExtReq q;
	q.ns = ns, q.as = as, q.ss = ss, q.nl = nl, q.al = al, q.opt = *opt, q.ok = 0, q.res = {};
	q.opt.flag &= ~(NS_F_CIGAR | NS_F_GPU);
	return q;
}
} // namespace

static int nsh_ext_one(const char *ns, int32_t nl, const char *as, int32_t al, const ns_opt_t *opt, const uint8_t *ss, ns_rst_t *r)
{
	if (!nsh_ensure_device()) return 0;
	ExtGen g;
	g.req.push_back(make_req(ns, nl, as, al, opt, ss));
	ext_exec().submit(&g);
	ext_exec().wait_until([&] { return g.done != 0; });
	if (!g.req[0].ok) return 0;
	r->n_cigar = 0, r->nt_len = g.req[0].res.nt_len, r->aa_len = g.req[0].res.aa_len, r->score = g.req[0].res.score;
	++g_dispatch;
	return 1;
}

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
	b->pending.push_back(make_req(ns, nl, as, al, opt, ss));
	return b->next_id++;
}

static void batch_reap(ns_ext_batch_t *b)
{
	std::vector<ExtGen*> keep;
	for (ExtGen *g : b->gens) {
		if (ext_exec().is_done(g) && g->n_got >= (int32_t)g->req.size()) delete g;
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
	g->base = b->next_id - n;
	g->req.swap(b->pending);
	b->gens.push_back(g);
	g_ext_req += n;
	g_dispatch += n;
	ext_exec().submit(g);
	return n;
}

static ExtGen *batch_find(ns_ext_batch_t *b, int32_t id)
{
	for (ExtGen *g : b->gens)
		if (id >= g->base && id < g->base + (int32_t)g->req.size()) return g;
	return nullptr;
}

extern "C" int ns_ext_batch_ready(ns_ext_batch_t *b, int32_t id)
{
	ExtGen *g = batch_find(b, id);
	return g && ext_exec().is_done(g);
}

extern "C" void ns_ext_batch_wait(ns_ext_batch_t *b)
{
	if (!b) return;
	// block until some generation of this batch that is still running completes (or none is)
	std::vector<ExtGen*> open;
	for (ExtGen *g : b->gens) if (!ext_exec().is_done(g)) open.push_back(g);
	if (open.empty()) return;
	ext_exec().wait_until([&] { for (ExtGen *g : open) if (g->done) return true; return false; });
}

extern "C" const ns_rst_t *ns_ext_batch_get(ns_ext_batch_t *b, int32_t id)
{
	ExtGen *g = batch_find(b, id);
	if (!g) return nullptr;
	ext_exec().wait_until([&] { return g->done != 0; });
	++g->n_got;
	return &g->rst[id - g->base];
}

extern "C" void ns_ext_batch_destroy(ns_ext_batch_t *b)
{
	if (!b) return;
	for (ExtGen *g : b->gens) {   // the executor may still be working on a generation
		ext_exec().wait_until([&] { return g->done != 0; });
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
