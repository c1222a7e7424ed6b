/* Request server: ECP-matrix requests of different molecules arrive at any
 * time (libecp_server_submit) and are evaluated with continuous batching.
 *
 * Each of `slots` worker threads owns a device context (its own HIP streams,
 * pinned staging and the resident static tables: grids, Bessel table,
 * angular and transformation tables).  A free worker takes every request
 * queued at that moment, up to a nominal-triplet cap, screens them into one
 * job on the host, and asks the shared device pool -- a single allocation of
 * the HBM budget -- for the job's device block.  The job is admitted when the
 * block fits; it runs while other workers' jobs run, so work that arrives
 * while the device is busy starts as soon as a slot and memory are free.
 * Results are scattered on the host into a packed upper triangle per
 * request and kept until the caller collects them.
 *
 * Every triplet is evaluated by the same kernels whatever else is in its
 * job, with the per-primitive-pair sums in a fixed order, so a request's
 * matrix does not depend on what it was batched with.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/libecp_ext.h"
#include "ecp_host.h"

enum { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2 };

typedef struct req {
  long long id;
  ecp_mol mol;
  int nao, nsh;
  long long nominal;
  double t_submit, t_admit, t_done;
  int state, rc;
  char err[192];
  double *upper; /* packed upper triangle, nao (nao + 1) / 2 */
  struct req *qnext, *dnext;
} req;

/* first-fit allocator over one device block, free ranges sorted by offset */
typedef struct {
  size_t off, size;
} range;

typedef struct {
  char *base;
  size_t cap, used, peak;
  size_t lost; /* bytes kept out of the pool after a failure */
  range *fr;
  int nfr, capfr;
} pool;

static size_t pool_take(pool *p, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  for (int i = 0; i < p->nfr; i++)
    if (p->fr[i].size >= bytes) {
      const size_t o = p->fr[i].off;
      p->fr[i].off += bytes;
      p->fr[i].size -= bytes;
      if (p->fr[i].size == 0) {
        memmove(p->fr + i, p->fr + i + 1, sizeof(range) * (size_t)(p->nfr - i - 1));
        p->nfr--;
      }
      p->used += bytes;
      if (p->used > p->peak) p->peak = p->used;
      return o;
    }
  return (size_t)-1;
}

/* Returns the range to the free list: 0, or -1 when the free list cannot
   grow (host memory); the range then stays out of the pool. */
static int pool_give(pool *p, size_t off, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  int i = 0;
  while (i < p->nfr && p->fr[i].off < off) i++;
  const int joinl = i > 0 && p->fr[i - 1].off + p->fr[i - 1].size == off;
  const int joinr = i < p->nfr && off + bytes == p->fr[i].off;
  if (!joinl && !joinr && p->nfr == p->capfr) {
    const int nc = p->capfr ? 2 * p->capfr : 16;
    range *nf = realloc(p->fr, sizeof(range) * (size_t)nc);
    if (!nf) return -1;
    p->fr = nf;
    p->capfr = nc;
  }
  p->used -= bytes;
  if (joinl && joinr) {
    p->fr[i - 1].size += bytes + p->fr[i].size;
    memmove(p->fr + i, p->fr + i + 1, sizeof(range) * (size_t)(p->nfr - i - 1));
    p->nfr--;
  } else if (joinl) {
    p->fr[i - 1].size += bytes;
  } else if (joinr) {
    p->fr[i].off = off;
    p->fr[i].size += bytes;
  } else {
    memmove(p->fr + i + 1, p->fr + i, sizeof(range) * (size_t)(p->nfr - i));
    p->fr[i].off = off;
    p->fr[i].size = bytes;
    p->nfr++;
  }
  return 0;
}

struct libecp_server;

typedef struct {
  struct libecp_server *srv;
  int slot;
  ecp_dev *dev;
  double *blocks;
  size_t cap_blocks;
  unsigned char *status;
  size_t cap_status;
  req **rs;     /* the requests of the current job */
  ecp_mol *mols;
  int cap_rs;
} worker;

struct libecp_server {
  pthread_mutex_t mu;
  pthread_cond_t work_cv, mem_cv, done_cv;
  int nslots, device, stop;
  pthread_t *th;
  worker *wk;
  ecp_static st;
  pool mem;
  long long max_batch_nominal;
  int max_batch_requests;
  req *qhead, *qtail;
  req *dhead, *dtail;
  req **tab;
  long long ntab, captab;
  double t0;
  /* statistics */
  long long submitted, completed, jobs, batched_requests;
  int running, max_running, max_batch_seen, queued;
};

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void finish(struct libecp_server *s, req *r, int rc, const char *msg) {
  /* caller holds s->mu */
  r->rc = rc;
  if (msg) snprintf(r->err, sizeof(r->err), "%s", msg);
  r->state = RQ_DONE;
  r->t_done = now_s() - s->t0;
  r->dnext = NULL;
  if (s->dtail)
    s->dtail->dnext = r;
  else
    s->dhead = r;
  s->dtail = r;
  s->completed++;
}

static void requeue_front(struct libecp_server *s, req **rs, int n) {
  /* caller holds s->mu; rs[0..n-1] go back to the queue head in order */
  for (int i = n - 1; i >= 0; i--) {
    rs[i]->state = RQ_QUEUED;
    rs[i]->qnext = s->qhead;
    s->qhead = rs[i];
    if (!s->qtail) s->qtail = rs[i];
    s->queued++;
  }
}

static void *worker_main(void *arg) {
  worker *w = arg;
  struct libecp_server *s = w->srv;
  char err[512];
  (void)ecp_dev_use(s->device);
  for (;;) {
    req **rs = w->rs;
    ecp_mol *mols = w->mols;
    /* take every queued request up to the batch cap */
    pthread_mutex_lock(&s->mu);
    while (!s->stop && !s->qhead) pthread_cond_wait(&s->work_cv, &s->mu);
    if (s->stop) {
      pthread_mutex_unlock(&s->mu);
      break;
    }
    int n = 0;
    long long nom = 0;
    while (s->qhead && n < s->max_batch_requests &&
           (n == 0 || nom + s->qhead->nominal <= s->max_batch_nominal)) {
      if (n == w->cap_rs) {
        /* grow the job's request arrays; without host memory the job keeps
           the requests taken so far */
        const int nc = 2 * w->cap_rs;
        req **nr = realloc(w->rs, sizeof(req *) * (size_t)nc);
        if (!nr) break;
        w->rs = rs = nr;
        ecp_mol *nm = realloc(w->mols, sizeof(ecp_mol) * (size_t)nc);
        if (!nm) break;
        w->mols = mols = nm;
        w->cap_rs = nc;
      }
      req *r = s->qhead;
      s->qhead = r->qnext;
      if (!s->qhead) s->qtail = NULL;
      s->queued--;
      r->state = RQ_RUNNING;
      rs[n++] = r;
      nom += r->nominal;
    }
    if (s->qhead) pthread_cond_signal(&s->work_cv);
    pthread_mutex_unlock(&s->mu);

    /* screen on the host and plan the job; a batch that does not fit the
       pool (or fails to screen) is split: the first request is retried
       alone and the rest go back to the queue */
    ecp_job job;
    size_t bytes = 0;
    int ok = 0;
    while (!ok) {
      for (int i = 0; i < n; i++) mols[i] = rs[i]->mol;
      err[0] = 0;
      if (ecp_job_build(mols, n, 0, &s->st, &job, err, sizeof(err))) {
        bytes = (size_t)-1;
      } else if (ecp_dev_plan(w->dev, &job, &bytes, err, sizeof(err))) {
        ecp_job_free(&job);
        bytes = (size_t)-1;
      } else if (bytes > s->mem.cap) {
        snprintf(err, sizeof(err), "request needs %zu bytes of device memory; the HBM budget is %zu", bytes,
                 s->mem.cap);
        ecp_job_free(&job);
        bytes = (size_t)-1;
      } else {
        ok = 1;
        break;
      }
      pthread_mutex_lock(&s->mu);
      if (n > 1) {
        requeue_front(s, rs + 1, n - 1);
        pthread_cond_signal(&s->work_cv);
        n = 1;
      } else {
        finish(s, rs[0], -1, err);
        pthread_cond_broadcast(&s->done_cv);
        n = 0;
      }
      pthread_mutex_unlock(&s->mu);
      if (n == 0) break;
    }
    if (!ok) continue;

    /* admission: wait for the job's block in the pool */
    pthread_mutex_lock(&s->mu);
    size_t off = (size_t)-1;
    if (bytes > 0)
      while ((off = pool_take(&s->mem, bytes)) == (size_t)-1) pthread_cond_wait(&s->mem_cv, &s->mu);
    const double t_admit = now_s() - s->t0;
    for (int i = 0; i < n; i++) rs[i]->t_admit = t_admit;
    s->running++;
    if (s->running > s->max_running) s->max_running = s->running;
    s->jobs++;
    s->batched_requests += n;
    if (n > s->max_batch_seen) s->max_batch_seen = n;
    pthread_mutex_unlock(&s->mu);

    /* evaluate */
    int dev_rc = 0;
    if (job.ntrip > 0) {
      if (job.nblk + 1 > w->cap_blocks) {
        free(w->blocks);
        w->cap_blocks = job.nblk + job.nblk / 4 + 1;
        w->blocks = malloc(sizeof(double) * w->cap_blocks);
        if (!w->blocks) w->cap_blocks = 0;
      }
      if (2 * (size_t)job.ntrip + 1 > w->cap_status) {
        free(w->status);
        w->cap_status = 2 * (size_t)job.ntrip + (size_t)job.ntrip / 2 + 1;
        w->status = malloc(w->cap_status);
        if (!w->status) w->cap_status = 0;
      }
      if (!w->blocks || !w->status) {
        snprintf(err, sizeof(err), "out of host memory for the blocks");
        dev_rc = -1;
      } else if (ecp_dev_bind(w->dev, s->mem.base + off, bytes, err, sizeof(err)) ||
          ecp_dev_execute(w->dev, 1, NULL, err, sizeof(err)) ||
          ecp_dev_fetch(w->dev, w->blocks, w->status, err, sizeof(err)))
        dev_rc = -1;
    }
    /* after a device error, kernels of the job may still be queued on the
       context's streams: wait for them before the job's memory is reused;
       if that fails too, the range stays out of the pool */
    const int reuse = !dev_rc || ecp_dev_sync(w->dev) == 0;
    pthread_mutex_lock(&s->mu);
    if (bytes > 0 && reuse && pool_give(&s->mem, off, bytes)) s->mem.lost += (bytes + 255) & ~(size_t)255;
    s->running--;
    pthread_cond_broadcast(&s->mem_cv);
    pthread_mutex_unlock(&s->mu);

    /* per-request matrices (host) */
    int *aoff = malloc(sizeof(int) * ((size_t)job.nsh + 1));
    if (!aoff && !dev_rc) {
      snprintf(err, sizeof(err), "out of host memory");
      dev_rc = -1;
    }
    if (aoff) ecp_job_ao_offsets(&job, aoff);
    int rcs[n > 0 ? n : 1];
    for (int i = 0; i < n; i++) {
      req *r = rs[i];
      rcs[i] = dev_rc;
      if (dev_rc) continue;
      const size_t np = (size_t)r->nao * (size_t)(r->nao + 1) / 2;
      r->upper = calloc(np ? np : 1, sizeof(double));
      if (!r->upper) {
        rcs[i] = -1;
        continue;
      }
      rcs[i] = job.ntrip > 0 ? ecp_job_scatter(&job, i, aoff, w->blocks, w->status, 0, r->nao, r->upper) : 0;
    }
    free(aoff);
    ecp_job_free(&job);
    pthread_mutex_lock(&s->mu);
    for (int i = 0; i < n; i++) finish(s, rs[i], rcs[i], rcs[i] < 0 ? (dev_rc ? err : "out of host memory") : NULL);
    pthread_cond_broadcast(&s->done_cv);
    pthread_mutex_unlock(&s->mu);
  }
  return NULL;
}

static void server_release(struct libecp_server *s) {
  for (int i = 0; i < s->nslots; i++) {
    ecp_dev_destroy(s->wk[i].dev);
    free(s->wk[i].blocks);
    free(s->wk[i].status);
    free(s->wk[i].rs);
    free(s->wk[i].mols);
  }
  for (long long i = 0; i < s->ntab; i++)
    if (s->tab[i]) {
      ecp_mol_release(&s->tab[i]->mol);
      free(s->tab[i]->upper);
      free(s->tab[i]);
    }
  ecp_dev_mem_free(s->mem.base);
  free(s->mem.fr);
  ecp_static_free(&s->st);
  free(s->tab);
  free(s->wk);
  free(s->th);
  pthread_mutex_destroy(&s->mu);
  pthread_cond_destroy(&s->work_cv);
  pthread_cond_destroy(&s->mem_cv);
  pthread_cond_destroy(&s->done_cv);
  free(s);
}

/* error text of the last failed call on this thread (ecp_api.c) */
void ecp_api_set_error(const char *msg);

libecp_server *libecp_server_create(int slots, size_t hbm_budget, long long max_batch_triplets,
                                    int largeGridOrder, double tolerance, double accuracy) {
  char err[512] = "";
  ecp_api_set_error("");
  if (slots < 1 || slots > 16 || hbm_budget < (1u << 20)) {
    ecp_api_set_error("libecp_server_create: slots must be 1..16 and the HBM budget at least 1 MiB");
    return NULL;
  }
  struct libecp_server *s = calloc(1, sizeof(*s));
  if (!s) {
    ecp_api_set_error("libecp_server_create: out of host memory");
    return NULL;
  }
  pthread_mutex_init(&s->mu, NULL);
  pthread_cond_init(&s->work_cv, NULL);
  pthread_cond_init(&s->mem_cv, NULL);
  pthread_cond_init(&s->done_cv, NULL);
  s->max_batch_nominal = max_batch_triplets > 0 ? max_batch_triplets : 4000000;
  s->max_batch_requests = 256;
  s->t0 = now_s();
  s->wk = calloc((size_t)slots, sizeof(worker));
  s->th = calloc((size_t)slots, sizeof(pthread_t));
  if (!s->wk || !s->th) {
    ecp_api_set_error("libecp_server_create: out of host memory");
    server_release(s);
    return NULL;
  }
  /* static tables sized for every request this build accepts */
  if (ecp_static_init(&s->st, ECP_MAX_L, ECP_MAX_LBS, NULL, -1, largeGridOrder, tolerance, accuracy, err,
                      sizeof(err))) {
    ecp_api_set_error(err);
    server_release(s);
    return NULL;
  }
  s->device = ecp_dev_current();
  for (int i = 0; i < slots; i++) {
    s->wk[i].srv = s;
    s->wk[i].slot = i;
    s->wk[i].cap_rs = 64;
    s->wk[i].rs = malloc(sizeof(req *) * 64);
    s->wk[i].mols = malloc(sizeof(ecp_mol) * 64);
    if (!s->wk[i].rs || !s->wk[i].mols) {
      s->nslots = i + 1;
      ecp_api_set_error("libecp_server_create: out of host memory");
      server_release(s);
      return NULL;
    }
    s->wk[i].dev = ecp_dev_create(&s->st, err, sizeof(err));
    if (!s->wk[i].dev) {
      s->nslots = i;
      ecp_api_set_error(err);
      server_release(s);
      return NULL;
    }
    s->nslots = i + 1;
  }
  void *base = NULL;
  if (ecp_dev_mem_alloc(&base, hbm_budget, err, sizeof(err))) {
    ecp_api_set_error(err);
    server_release(s);
    return NULL;
  }
  s->mem.base = base;
  s->mem.cap = hbm_budget & ~(size_t)255;
  s->mem.capfr = 16;
  s->mem.fr = malloc(sizeof(range) * 16);
  if (!s->mem.fr) {
    ecp_api_set_error("libecp_server_create: out of host memory");
    server_release(s);
    return NULL;
  }
  s->mem.fr[0].off = 0;
  s->mem.fr[0].size = s->mem.cap;
  s->mem.nfr = 1;
  for (int i = 0; i < slots; i++)
    if (pthread_create(&s->th[i], NULL, worker_main, &s->wk[i])) {
      pthread_mutex_lock(&s->mu);
      s->stop = 1;
      pthread_cond_broadcast(&s->work_cv);
      pthread_mutex_unlock(&s->mu);
      for (int j = 0; j < i; j++) pthread_join(s->th[j], NULL);
      ecp_api_set_error("pthread_create failed");
      server_release(s);
      return NULL;
    }
  return s;
}

long long libecp_server_submit(libecp_server *s, const libecp_molecule *M) {
  char err[512] = "";
  if (!s || !M) return -1;
  req *r = calloc(1, sizeof(req));
  if (!r) {
    ecp_api_set_error("libecp_server_submit: out of host memory");
    return -1;
  }
  ecp_mol m0 = {M->nrAtoms, M->geometry, M->shellsECP, M->lECP, M->KECP, M->nECP, M->dECP, M->aECP,
                M->shellsBS, M->lBS, M->KBS, M->dBS, M->aBS};
  ecp_mol_dims d;
  if (ecp_mol_copy(&m0, &r->mol, &d, err, sizeof(err))) {
    ecp_api_set_error(err);
    free(r);
    return -1;
  }
  r->nao = d.nao;
  r->nsh = d.nsh;
  r->nominal = (long long)d.necp * d.nsh * (d.nsh + 1) / 2;
  pthread_mutex_lock(&s->mu);
  if (s->ntab == s->captab) {
    const long long nc = s->captab ? 2 * s->captab : 1024;
    req **nt = realloc(s->tab, sizeof(req *) * (size_t)nc);
    if (!nt) {
      pthread_mutex_unlock(&s->mu);
      ecp_mol_release(&r->mol);
      free(r);
      ecp_api_set_error("libecp_server_submit: out of host memory");
      return -1;
    }
    s->tab = nt;
    s->captab = nc;
  }
  r->id = s->ntab;
  r->t_submit = now_s() - s->t0;
  s->tab[s->ntab++] = r;
  r->qnext = NULL;
  if (s->qtail)
    s->qtail->qnext = r;
  else
    s->qhead = r;
  s->qtail = r;
  s->queued++;
  s->submitted++;
  pthread_cond_signal(&s->work_cv);
  pthread_mutex_unlock(&s->mu);
  return r->id;
}

static req *lookup(libecp_server *s, long long t) { return (t >= 0 && t < s->ntab) ? s->tab[t] : NULL; }

static int wait_until(pthread_cond_t *cv, pthread_mutex_t *mu, double deadline) {
  struct timespec ts;
  clock_gettime(CLOCK_REALTIME, &ts);
  const double rel = deadline - now_s();
  if (rel <= 0) return 1;
  double sec = (double)ts.tv_sec + 1e-9 * ts.tv_nsec + rel;
  ts.tv_sec = (time_t)sec;
  ts.tv_nsec = (long)((sec - (double)ts.tv_sec) * 1e9);
  return pthread_cond_timedwait(cv, mu, &ts) != 0 && now_s() >= deadline;
}

int libecp_server_wait(libecp_server *s, long long ticket, double timeout_s) {
  if (!s) return -1;
  const double deadline = now_s() + (timeout_s < 0 ? 1e9 : timeout_s);
  pthread_mutex_lock(&s->mu);
  req *r = lookup(s, ticket);
  int rc = -1;
  if (r) {
    while (r->state != RQ_DONE)
      if (wait_until(&s->done_cv, &s->mu, deadline)) break;
    rc = r->state == RQ_DONE ? 0 : 1;
  }
  pthread_mutex_unlock(&s->mu);
  return rc;
}

long long libecp_server_next(libecp_server *s, double timeout_s) {
  if (!s) return -1;
  const double deadline = now_s() + (timeout_s < 0 ? 1e9 : timeout_s);
  pthread_mutex_lock(&s->mu);
  while (!s->dhead)
    if (wait_until(&s->done_cv, &s->mu, deadline)) break;
  long long t = -1;
  if (s->dhead) {
    req *r = s->dhead;
    s->dhead = r->dnext;
    if (!s->dhead) s->dtail = NULL;
    t = r->id;
  }
  pthread_mutex_unlock(&s->mu);
  return t;
}

int libecp_server_result(libecp_server *s, long long ticket, int rowdim, double *I, double *times) {
  if (!s) return -1;
  pthread_mutex_lock(&s->mu);
  req *r = lookup(s, ticket);
  if (!r || r->state != RQ_DONE) {
    pthread_mutex_unlock(&s->mu);
    ecp_api_set_error(r ? "request not finished" : "unknown or already collected ticket");
    return -1;
  }
  s->tab[ticket] = NULL;
  /* drop it from the completion list if it is still there */
  req **pp = &s->dhead, *prev = NULL;
  while (*pp && *pp != r) {
    prev = *pp;
    pp = &(*pp)->dnext;
  }
  if (*pp) {
    *pp = r->dnext;
    if (s->dtail == r) s->dtail = prev;
  }
  pthread_mutex_unlock(&s->mu);
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_admit;
    times[2] = r->t_done;
  }
  const int rc = r->rc;
  if (rc < 0)
    ecp_api_set_error(r->err);
  else if (I && r->upper)
    for (int i = 0; i < r->nao; i++) {
      const double *row = r->upper + (size_t)i * r->nao - (size_t)i * (i - 1) / 2 - i;
      for (int j = i; j < r->nao; j++) I[(size_t)i * rowdim + j] += row[j];
    }
  ecp_mol_release(&r->mol);
  free(r->upper);
  free(r);
  return rc;
}

int libecp_server_stats(libecp_server *s, long long *out, int n) {
  if (!s) return -1;
  pthread_mutex_lock(&s->mu);
  const long long v[] = {s->submitted,        s->completed,   s->jobs,
                         s->batched_requests, s->max_running, s->max_batch_seen,
                         (long long)s->mem.peak, (long long)s->mem.cap, s->queued,
                         s->running};
  pthread_mutex_unlock(&s->mu);
  const int k = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < k && i < n; i++) out[i] = v[i];
  return k;
}

void libecp_server_free(libecp_server *s) {
  if (!s) return;
  pthread_mutex_lock(&s->mu);
  s->stop = 1;
  pthread_cond_broadcast(&s->work_cv);
  pthread_cond_broadcast(&s->mem_cv);
  pthread_mutex_unlock(&s->mu);
  for (int i = 0; i < s->nslots; i++) pthread_join(s->th[i], NULL);
  server_release(s);
}
