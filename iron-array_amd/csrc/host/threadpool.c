/*
 * threadpool.c -- a minimal parallel-for for the host stages of libiarray.so
 * (chunk encoding).  The reference parallelises the same work over Blosc2
 * blocks with cfg.max_num_threads threads (src/iarray.c:406); the host encode
 * here uses every core the process may run on unless IARRAY_HOST_THREADS says
 * otherwise.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <time.h>

#include "iarray_private.h"

double iax_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int iax_nthreads(const iarray_config_t *cfg) {
  const char *e = getenv("IARRAY_HOST_THREADS");
  if (e && atoi(e) > 0) return atoi(e);
  cpu_set_t set;
  int n = 0;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = 1;
  if (n > 64) n = 64;
  (void)cfg;
  return n;
}

typedef struct {
  iax_task_fn fn;
  void *ctx;
  int64_t n;
  atomic_llong next;
} job_t;

typedef struct {
  job_t *job;
  int worker;
} arg_t;

static void *run(void *p) {
  arg_t *a = (arg_t *)p;
  for (;;) {
    const int64_t i = atomic_fetch_add(&a->job->next, 1);
    if (i >= a->job->n) break;
    a->job->fn(a->job->ctx, i, a->worker);
  }
  return NULL;
}

void iax_parallel_for(int nthreads, int64_t n, iax_task_fn fn, void *ctx) {
  job_t job;
  job.fn = fn;
  job.ctx = ctx;
  job.n = n;
  atomic_init(&job.next, 0);
  if (nthreads > n) nthreads = (int)n;
  if (nthreads <= 1) {
    arg_t a = {&job, 0};
    run(&a);
    return;
  }
  pthread_t th[64];
  arg_t args[64];
  int started = 0;
  for (int t = 1; t < nthreads && t < 64; t++) {
    args[t].job = &job;
    args[t].worker = t;
    if (pthread_create(&th[t], NULL, run, &args[t]) == 0) started = t;
    else break;
  }
  args[0].job = &job;
  args[0].worker = 0;
  run(&args[0]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
}
