/* Drives the library's reference entry points the way the reference's callers do, from plain C
 * threads that take work from a shared counter as kt_for does (kthread.c:54-72), so that the
 * benchmark times the drop-in path without interpreter overhead:
 *   nshd_ext_batch   per group of `per_batch` requests (one query's candidates): ns_ext_batch_init,
 *                    ns_ext_batch_add of each, ns_ext_batch_flush, ns_ext_batch_ready polling with
 *                    ns_ext_batch_wait, ns_ext_batch_get, ns_ext_batch_destroy (mp_align_batch,
 *                    align.c:745-777);
 *   nshd_ext_single  one ns_global_gs16b call per request (the synchronous mp_align_ext_issue path,
 *                    align.c:371-381);
 *   nshd_gpu_submit  per group: ns_gpu_submit of every request, then ns_gpu_wait of each
 *                    (mp_align_gap's submit-all-then-wait, align.c:437-466).
 * Results land in out (3 ints per request for extensions: nt_len, aa_len, score; for the global
 * DP: score and CIGAR length); *wall receives the elapsed seconds. */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>
#include "../../include/nasw-hip.h"

typedef struct {
	int32_t n, per_batch, kind;
	const char *ns; const int64_t *ns_off; const int32_t *nl;
	const char *as; const int64_t *as_off; const int32_t *al;
	const uint8_t *dirs; const ns_opt_t *opt;
	int32_t *out;
	volatile int32_t next;
	volatile int32_t failed;
} job_t;

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void opt_of(const job_t *j, int32_t k, ns_opt_t *o)
{
	*o = *j->opt;
	if (j->dirs) o->flag = (o->flag & ~(NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) | (j->dirs[k]? NS_F_EXT_LEFT : NS_F_EXT_RIGHT);
}

static void *worker(void *arg)
{
	job_t *j = (job_t*)arg;
	ns_opt_t *opts = (ns_opt_t*)malloc(sizeof(ns_opt_t) * j->per_batch);
	int32_t *ids = (int32_t*)malloc(sizeof(int32_t) * j->per_batch);
	ns_gpu_request_t **qs = (ns_gpu_request_t**)malloc(sizeof(void*) * j->per_batch);
	for (;;) {
		int32_t k0 = __sync_fetch_and_add(&j->next, j->per_batch), k1, k;
		if (k0 >= j->n) break;
		k1 = k0 + j->per_batch < j->n? k0 + j->per_batch : j->n;
		if (j->kind == 0) {
			ns_ext_batch_t *b = ns_ext_batch_init(0);
			int32_t pending;
			if (!b) { j->failed = 1; break; }
			for (k = k0; k < k1; ++k) {
				opt_of(j, k, &opts[k - k0]);
				ids[k - k0] = ns_ext_batch_add(b, j->ns + j->ns_off[k], j->nl[k], j->as + j->as_off[k], j->al[k], &opts[k - k0], 0);
			}
			ns_ext_batch_flush(b);
			do {   // poll as mp_align_batch does, waiting when nothing is ready
				pending = 0;
				for (k = k0; k < k1; ++k) pending += !ns_ext_batch_ready(b, ids[k - k0]);
				if (pending) ns_ext_batch_wait(b);
			} while (pending);
			for (k = k0; k < k1; ++k) {
				const ns_rst_t *r = ns_ext_batch_get(b, ids[k - k0]);
				j->out[3 * k] = r->nt_len, j->out[3 * k + 1] = r->aa_len, j->out[3 * k + 2] = r->score;
			}
			ns_ext_batch_destroy(b);
		} else if (j->kind == 1) {
			for (k = k0; k < k1; ++k) {
				ns_rst_t r;
				ns_opt_t o;
				opt_of(j, k, &o);
				ns_rst_init(&r);
				ns_global_gs16b(0, j->ns + j->ns_off[k], j->nl[k], j->as + j->as_off[k], j->al[k], &o, 0, &r);
				j->out[3 * k] = r.nt_len, j->out[3 * k + 1] = r.aa_len, j->out[3 * k + 2] = r.score;
				free(r.cigar);
			}
		} else {
			for (k = k0; k < k1; ++k)
				qs[k - k0] = ns_gpu_submit(0, j->ns + j->ns_off[k], j->nl[k], j->as + j->as_off[k], j->al[k], j->opt, 0);
			for (k = k0; k < k1; ++k) {
				ns_rst_t r;
				ns_rst_init(&r);
				if (qs[k - k0] && ns_gpu_wait(qs[k - k0], &r)) j->out[3 * k] = r.score, j->out[3 * k + 1] = r.n_cigar;
				else j->out[3 * k] = INT32_MIN, j->out[3 * k + 1] = -1, j->failed = 1;
				j->out[3 * k + 2] = 0;
				free(r.cigar);
			}
		}
	}
	free(opts), free(ids), free(qs);
	return 0;
}

static int32_t run(int32_t kind, int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt, int32_t n_threads,
		int32_t per_batch, int32_t *out, double *wall)
{
	job_t j;
	pthread_t *tid;
	int32_t t;
	double t0;
	if (n_threads < 1) n_threads = 1;
	if (per_batch < 1) per_batch = 1;
	j.n = n, j.per_batch = per_batch, j.kind = kind, j.ns = ns, j.ns_off = ns_off, j.nl = nl, j.as = as;
	j.as_off = as_off, j.al = al, j.dirs = dirs, j.opt = opt, j.out = out, j.next = 0, j.failed = 0;
	tid = (pthread_t*)malloc(sizeof(pthread_t) * n_threads);
	t0 = now_s();
	for (t = 0; t < n_threads; ++t) pthread_create(&tid[t], 0, worker, &j);
	for (t = 0; t < n_threads; ++t) pthread_join(tid[t], 0);
	*wall = now_s() - t0;
	free(tid);
	return !j.failed;
}

int32_t nshd_ext_batch(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt, int32_t n_threads,
		int32_t per_batch, int32_t *out, double *wall)
{
	return run(0, n, ns, ns_off, nl, as, as_off, al, dirs, opt, n_threads, per_batch, out, wall);
}

int32_t nshd_ext_single(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt, int32_t n_threads,
		int32_t *out, double *wall)
{
	return run(1, n, ns, ns_off, nl, as, as_off, al, dirs, opt, n_threads, 64, out, wall);
}

int32_t nshd_gpu_submit(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const ns_opt_t *opt, int32_t n_threads, int32_t per_batch,
		int32_t *out, double *wall)
{
	return run(2, n, ns, ns_off, nl, as, as_off, al, 0, opt, n_threads, per_batch, out, wall);
}
