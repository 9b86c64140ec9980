/* Test infrastructure, built only into oracle/_ref/libnsref.so by `make ref` when the
 * reference tree is present: the reference's score-only extension DP (ns_global_gs16b,
 * nasw-sse.c:763-772) over a batch of requests from plain pthreads, each thread taking the next
 * request from a shared counter as kt_for does (kthread.c:54-72). It is the CPU baseline of the
 * batched extension benchmark and the source of the reference results that the GPU results are
 * compared with. Nothing here is part of the product. */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>
#include "nasw.h"

typedef struct {
	int32_t n;
	const char *ns; const int64_t *ns_off; const int32_t *nl;
	const char *as; const int64_t *as_off; const int32_t *al;
	const uint8_t *dirs;         /* 1: EXT_LEFT, else EXT_RIGHT; NULL: opt->flag as given */
	const ns_opt_t *opt;
	double deadline;             /* seconds since the start; <= 0: no limit */
	double t0;
	int32_t *out;                /* 3 ints per request: nt_len, aa_len, score */
	volatile int32_t next, done;
} bench_job_t;

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void *bench_worker(void *arg)
{
	bench_job_t *j = (bench_job_t*)arg;
	for (;;) {
		int32_t k = __sync_fetch_and_add(&j->next, 1);
		ns_opt_t o;
		ns_rst_t r;
		if (k >= j->n) break;
		if (j->deadline > 0 && now_s() - j->t0 > j->deadline) break;
		o = *j->opt;
		if (j->dirs) o.flag = (o.flag & ~(NS_F_EXT_LEFT | NS_F_EXT_RIGHT)) | (j->dirs[k]? NS_F_EXT_LEFT : NS_F_EXT_RIGHT);
		ns_rst_init(&r);
		ns_global_gs16b(0, j->ns + j->ns_off[k], j->nl[k], j->as + j->as_off[k], j->al[k], &o, 0, &r);
		j->out[3 * k] = r.nt_len, j->out[3 * k + 1] = r.aa_len, j->out[3 * k + 2] = r.score;
		free(r.cigar);
		__sync_fetch_and_add(&j->done, 1);
	}
	return 0;
}

/* Runs requests in index order over n_threads threads until all are done or `seconds` have
 * passed (<= 0: no limit). Requests started before the limit finish. Returns the number done
 * (a prefix of the order, up to the requests in flight when the limit passed: out[] of a request
 * not done is left as it was); *wall receives the elapsed seconds. */
int32_t or_ref_ext_batch(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const uint8_t *dirs, const ns_opt_t *opt, int32_t n_threads,
		double seconds, int32_t *out, double *wall)
{
	bench_job_t j;
	pthread_t *tid;
	int32_t t;
	if (n_threads < 1) n_threads = 1;
	j.n = n, j.ns = ns, j.ns_off = ns_off, j.nl = nl, j.as = as, j.as_off = as_off, j.al = al, j.dirs = dirs;
	j.opt = opt, j.deadline = seconds, j.out = out, j.next = 0, j.done = 0;
	tid = (pthread_t*)malloc(sizeof(pthread_t) * n_threads);
	j.t0 = now_s();
	for (t = 0; t < n_threads; ++t) pthread_create(&tid[t], 0, bench_worker, &j);
	for (t = 0; t < n_threads; ++t) pthread_join(tid[t], 0);
	*wall = now_s() - j.t0;
	free(tid);
	return j.done;
}

typedef struct {
	int32_t n;
	const char *ns; const int64_t *ns_off; const int32_t *nl;
	const char *as; const int64_t *as_off; const int32_t *al;
	const ns_opt_t *opt;
	int32_t *out;                /* 2 ints per request: score, CIGAR length */
	volatile int32_t next;
} glob_job_t;

static void *glob_worker(void *arg)
{
	glob_job_t *j = (glob_job_t*)arg;
	for (;;) {
		int32_t k = __sync_fetch_and_add(&j->next, 1);
		ns_rst_t r;
		if (k >= j->n) break;
		ns_rst_init(&r);
		ns_global_gs16b(0, j->ns + j->ns_off[k], j->nl[k], j->as + j->as_off[k], j->al[k], j->opt, 0, &r);
		j->out[2 * k] = r.score, j->out[2 * k + 1] = r.n_cigar;
		free(r.cigar);
	}
	return 0;
}

/* The reference's global spliced DP with CIGAR (ns_global_gs16b with NS_F_CIGAR in opt->flag) over
 * a batch from n_threads threads: out[2k] = score, out[2k+1] = CIGAR length; *wall = seconds. */
void or_ref_global_batch(int32_t n, const char *ns, const int64_t *ns_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, const ns_opt_t *opt, int32_t n_threads, int32_t *out, double *wall)
{
	glob_job_t j;
	pthread_t *tid;
	int32_t t;
	double t0;
	if (n_threads < 1) n_threads = 1;
	j.n = n, j.ns = ns, j.ns_off = ns_off, j.nl = nl, j.as = as, j.as_off = as_off, j.al = al, j.opt = opt;
	j.out = out, j.next = 0;
	tid = (pthread_t*)malloc(sizeof(pthread_t) * n_threads);
	t0 = now_s();
	for (t = 0; t < n_threads; ++t) pthread_create(&tid[t], 0, glob_worker, &j);
	for (t = 0; t < n_threads; ++t) pthread_join(tid[t], 0);
	*wall = now_s() - t0;
	free(tid);
}
