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
#include <stdint.h>
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

/* Per-call threads: the fallback when the pool is busy (a second caller
 * thread) or cannot grow. */
static void spawn_for(job_t *job, int nthreads) {
  pthread_t th[64];
  arg_t args[64];
  int started = 0;
  for (int t = 1; t < nthreads && t < 64; t++) {
    args[t].job = job;
    args[t].worker = t;
    if (pthread_create(&th[t], NULL, run, &args[t]) == 0) started = t;
    else break;
  }
  args[0].job = job;
  args[0].worker = 0;
  run(&args[0]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
}

/* Persistent workers, started on first use and kept waiting on a condition
 * variable, so a parallel-for over a few chunks does not pay for thread
 * creation.  One job at a time: `want` workers join generation `gen` (worker
 * ids 1..want in joining order); the caller runs as worker 0 and returns once
 * all of them have joined and left, so no worker touches the job afterwards. */
static struct {
  pthread_mutex_t mu;
  pthread_cond_t work, done;
  int nworkers;
  unsigned long gen;
  job_t *job;
  int want, joined, active;
} g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, NULL, 0, 0, 0};
static pthread_mutex_t g_pool_owner = PTHREAD_MUTEX_INITIALIZER;
static pthread_once_t g_pool_once = PTHREAD_ONCE_INIT;

/* `start` is the generation current when the worker was created (before the
 * job that created it), so that job is the first it sees */
static void *pool_worker(void *start) {
  pthread_mutex_lock(&g_pool.mu);
  unsigned long seen = (unsigned long)(uintptr_t)start;
  for (;;) {
    while (g_pool.gen == seen) pthread_cond_wait(&g_pool.work, &g_pool.mu);
    seen = g_pool.gen;
    if (g_pool.joined >= g_pool.want) continue;
    arg_t a = {g_pool.job, ++g_pool.joined};
    g_pool.active++;
    pthread_mutex_unlock(&g_pool.mu);
    run(&a);
    pthread_mutex_lock(&g_pool.mu);
    if (--g_pool.active == 0 && g_pool.joined >= g_pool.want) pthread_cond_signal(&g_pool.done);
  }
  return NULL;
}

/* a forked child has none of the parent's workers */
static void pool_after_fork_child(void) {
  pthread_mutex_init(&g_pool.mu, NULL);
  pthread_cond_init(&g_pool.work, NULL);
  pthread_cond_init(&g_pool.done, NULL);
  pthread_mutex_init(&g_pool_owner, NULL);
  g_pool.nworkers = 0;
  g_pool.job = NULL;
  g_pool.want = g_pool.joined = g_pool.active = 0;
}

static void pool_init_once(void) { pthread_atfork(NULL, NULL, pool_after_fork_child); }

void iax_parallel_for(int nthreads, int64_t n, iax_task_fn fn, void *ctx) {
  job_t job;
  job.fn = fn;
  job.ctx = ctx;
  job.n = n;
  atomic_init(&job.next, 0);
  if (nthreads > n) nthreads = (int)n;
  if (nthreads > 64) nthreads = 64;
  if (nthreads <= 1) {
    arg_t a = {&job, 0};
    run(&a);
    return;
  }
  pthread_once(&g_pool_once, pool_init_once);
  if (pthread_mutex_trylock(&g_pool_owner) != 0) {
    spawn_for(&job, nthreads);
    return;
  }
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.nworkers < nthreads - 1) {
    pthread_t th;
    pthread_attr_t at;
    pthread_attr_init(&at);
    pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
    const int ok = pthread_create(&th, &at, pool_worker, (void *)(uintptr_t)g_pool.gen) == 0;
    pthread_attr_destroy(&at);
    if (!ok) break;
    g_pool.nworkers++;
  }
  g_pool.job = &job;
  g_pool.want = nthreads - 1 < g_pool.nworkers ? nthreads - 1 : g_pool.nworkers;
  g_pool.joined = 0;
  g_pool.active = 0;
  g_pool.gen++;
  pthread_cond_broadcast(&g_pool.work);
  pthread_mutex_unlock(&g_pool.mu);
  arg_t a0 = {&job, 0};
  run(&a0);
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.joined < g_pool.want || g_pool.active > 0) pthread_cond_wait(&g_pool.done, &g_pool.mu);
  g_pool.job = NULL;
  pthread_mutex_unlock(&g_pool.mu);
  pthread_mutex_unlock(&g_pool_owner);
}
