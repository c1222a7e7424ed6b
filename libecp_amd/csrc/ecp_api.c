/* Public C API of the MI355X ECP-integral library: the reference's libecp.h
 * handle interface and its one-call getIntegrals() (include/libecp.h,
 * include/getIntegrals.h), backed by the HIP pipeline in ecp_device.hip.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/libecp.h"
#include "../../include/getIntegrals.h"
#include "../../include/libecp_ext.h"
#include "ecp_host.h"

static __thread char g_err[512];

static void set_last_error(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg ? msg : ""); }

const char *libecp_last_error(void) { return g_err; }

void ecp_api_set_error(const char *msg) { set_last_error(msg); }

/* Phases of the last getIntegrals / calculateECPIntegrals call on this
   thread (ms): host tables, screening, device context, plan + upload,
   kernels, download, matrix / callbacks. */
enum { PH_TABLES = 0, PH_SCREEN, PH_CONTEXT, PH_UPLOAD, PH_KERNELS, PH_DOWNLOAD, PH_DELIVER, PH_N };
static __thread double g_phase[PH_N];

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

int libecp_last_call_ms(double *ms, int n) {
  for (int i = 0; i < PH_N && i < n; i++) ms[i] = g_phase[i];
  return PH_N;
}

/* ---- process-wide cache of static tables and device contexts -------------
   A handle needs, besides its molecule, geometry-independent tables (grids,
   Bessel table, angular and transformation tables: a few ms to build on the
   host) and a device context (streams, events, pinned staging, the tables
   resident in HBM).  Both depend only on (maxL, maxAlpha, grid order,
   tolerance, accuracy) and the default Cartesian ordering, so they are kept
   per process: the tables once per key, and up to kCacheIdle idle contexts
   per key and device, which a handle takes at its first calculation and
   returns when it is freed.  Handles with a custom ordering, and every
   handle while the cache is disabled, own their tables and context. */
enum { kCacheKeys = 8, kCacheIdle = 2 };
typedef struct {
  int used;
  int maxL, maxAlpha, grid;
  double tol, acc;
  ecp_static st; /* read-only once built */
  ecp_dev *idle[kCacheIdle];
  int idle_dev[kCacheIdle];
  int nidle;
} cache_slot;

static pthread_mutex_t g_cache_mu = PTHREAD_MUTEX_INITIALIZER;
static cache_slot g_cache[kCacheKeys];
static int g_cache_on = 1;

/* the slot of a key, built on first use; NULL when the cache is off or full
   (the caller then builds its own tables) */
static cache_slot *cache_tables(int maxL, int maxAlpha, int grid, double tol, double acc, char *err,
                                size_t errlen, int *failed) {
  *failed = 0;
  pthread_mutex_lock(&g_cache_mu);
  cache_slot *hit = NULL, *fresh = NULL;
  if (g_cache_on)
    for (int i = 0; i < kCacheKeys; i++) {
      cache_slot *c = &g_cache[i];
      if (!c->used) {
        if (!fresh) fresh = c;
        continue;
      }
      if (c->maxL == maxL && c->maxAlpha == maxAlpha && c->grid == grid && c->tol == tol && c->acc == acc) {
        hit = c;
        break;
      }
    }
  if (!hit && fresh) {
    if (ecp_static_init(&fresh->st, maxL, maxAlpha, NULL, -1, grid, tol, acc, err, errlen)) {
      *failed = 1;
    } else {
      fresh->used = 1;
      fresh->maxL = maxL;
      fresh->maxAlpha = maxAlpha;
      fresh->grid = grid;
      fresh->tol = tol;
      fresh->acc = acc;
      fresh->nidle = 0;
      hit = fresh;
    }
  }
  pthread_mutex_unlock(&g_cache_mu);
  return hit;
}

static ecp_dev *cache_take(cache_slot *c, int device) {
  ecp_dev *d = NULL;
  pthread_mutex_lock(&g_cache_mu);
  for (int i = c->nidle - 1; i >= 0; i--)
    if (c->idle_dev[i] == device) {
      d = c->idle[i];
      c->idle[i] = c->idle[c->nidle - 1];
      c->idle_dev[i] = c->idle_dev[c->nidle - 1];
      c->nidle--;
      break;
    }
  pthread_mutex_unlock(&g_cache_mu);
  return d;
}

/* keeps an idle context for the next handle of the same key, or frees it */
static void cache_give(cache_slot *c, ecp_dev *d, int device) {
  if (!d) return;
  pthread_mutex_lock(&g_cache_mu);
  if (g_cache_on && c->nidle < kCacheIdle) {
    c->idle[c->nidle] = d;
    c->idle_dev[c->nidle] = device;
    c->nidle++;
    d = NULL;
  }
  pthread_mutex_unlock(&g_cache_mu);
  ecp_dev_destroy(d);
}

void libecp_cache_clear(void) {
  ecp_dev *drop[kCacheKeys * kCacheIdle];
  int nd = 0;
  pthread_mutex_lock(&g_cache_mu);
  for (int i = 0; i < kCacheKeys; i++) {
    for (int k = 0; k < g_cache[i].nidle; k++) drop[nd++] = g_cache[i].idle[k];
    g_cache[i].nidle = 0;
  }
  pthread_mutex_unlock(&g_cache_mu);
  /* the tables stay: handles alive may read them; they are a few MB */
  for (int i = 0; i < nd; i++) ecp_dev_destroy(drop[i]);
}

int libecp_cache_enable(int on) {
  pthread_mutex_lock(&g_cache_mu);
  const int was = g_cache_on;
  g_cache_on = on != 0;
  pthread_mutex_unlock(&g_cache_mu);
  if (!on) libecp_cache_clear();
  return was;
}

struct _libECPHandle {
  int nat, n;
  /* private copies of the inputs (the reference borrows them, libecp.c:68-74) */
  double *geom;
  int *shellsECP, *lECP, *KECP;
  double *nECP, *dECP, *aECP;
  int *shellsBS, *lBS, *KBS;
  double *dBS, *aBS;
  ecp_mol mol;
  const ecp_static *st; /* &st_own, or the cache's tables for this key */
  ecp_static st_own;
  cache_slot *slot;     /* non-NULL: tables and context come from the cache */
  int dev_id;           /* HIP device of dev */
  ecp_dev *dev;
  ecp_job job;
  int job_ready;
  double *blocks;
  unsigned char *status;
  size_t cap_blocks, cap_status;
};

static void *dupmem(const void *src, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (p && bytes) memcpy(p, src, bytes);
  return p;
}

void libECP_free(libECPHandle *h) {
  if (!h) return;
  if (h->job_ready) ecp_job_free(&h->job);
  if (h->slot)
    cache_give(h->slot, h->dev, h->dev_id);
  else
    ecp_dev_destroy(h->dev);
  ecp_static_free(&h->st_own);
  free(h->geom);
  free(h->shellsECP);
  free(h->lECP);
  free(h->KECP);
  free(h->nECP);
  free(h->dECP);
  free(h->aECP);
  free(h->shellsBS);
  free(h->lBS);
  free(h->KBS);
  free(h->dBS);
  free(h->aBS);
  free(h->blocks);
  free(h->status);
  free(h);
}

libECPHandle *libECP_init(int nrAtoms, double *geometry, int *shellsECP, int *lECP, int *KECP, double *nECP,
                          double *dECP, double *aECP, int *shellsBS, int *lBS, int *KBS, double *dBS,
                          double *aBS, int n, int lmax, int *shellOrdering, int largeGridOrder, double tolerance,
                          double accuracy) {
  char err[512] = "";
  set_last_error("");
  if (n < 0 || n > 2) {
    set_last_error("derivative order n must be 0, 1 or 2");
    return NULL;
  }
  if (nrAtoms <= 0 || !geometry || !shellsECP || !shellsBS) {
    set_last_error("invalid arguments");
    return NULL;
  }
  libECPHandle *h = calloc(1, sizeof(*h));
  if (!h) {
    set_last_error("out of host memory");
    return NULL;
  }
  h->nat = nrAtoms;
  h->n = n;
  ecp_mol m0 = {nrAtoms, geometry, shellsECP, lECP, KECP, nECP, dECP, aECP, shellsBS, lBS, KBS, dBS, aBS};
  ecp_mol_dims d;
  if (ecp_mol_scan(&m0, &d, err, sizeof(err))) {
    set_last_error(err);
    free(h);
    return NULL;
  }
  if (d.necp == 0) {
    /* no ECP: the reference's tables would be built with maxLECP = 0; nothing
       is ever integrated.  Keep a valid handle that delivers no blocks. */
  }
  int necps = 0;
  for (int i = 0; i < nrAtoms; i++) necps += shellsECP[i];
  h->geom = dupmem(geometry, sizeof(double) * 3 * (size_t)nrAtoms);
  h->shellsECP = dupmem(shellsECP, sizeof(int) * (size_t)nrAtoms);
  h->lECP = dupmem(lECP, sizeof(int) * (size_t)necps);
  h->KECP = dupmem(KECP, sizeof(int) * (size_t)necps);
  h->nECP = dupmem(nECP, sizeof(double) * (size_t)d.ngauss);
  h->dECP = dupmem(dECP, sizeof(double) * (size_t)d.ngauss);
  h->aECP = dupmem(aECP, sizeof(double) * (size_t)d.ngauss);
  h->shellsBS = dupmem(shellsBS, sizeof(int) * (size_t)nrAtoms);
  h->lBS = dupmem(lBS, sizeof(int) * (size_t)d.nsh);
  h->KBS = dupmem(KBS, sizeof(int) * (size_t)d.nsh);
  h->dBS = dupmem(dBS, sizeof(double) * (size_t)d.nprim);
  h->aBS = dupmem(aBS, sizeof(double) * (size_t)d.nprim);
  ecp_mol m = {nrAtoms, h->geom, h->shellsECP, h->lECP, h->KECP, h->nECP, h->dECP, h->aECP,
               h->shellsBS, h->lBS, h->KBS, h->dBS, h->aBS};
  h->mol = m;
  const int maxL = d.maxLECP > 0 ? d.maxLECP : 1;
  const double t0 = now_ms();
  int failed = 0;
  if (!shellOrdering)
    h->slot = cache_tables(maxL, d.maxLBS + n, largeGridOrder, tolerance, accuracy, err, sizeof(err), &failed);
  if (h->slot) {
    h->st = &h->slot->st;
  } else if (failed || ecp_static_init(&h->st_own, maxL, d.maxLBS + n, shellOrdering, lmax, largeGridOrder,
                                       tolerance, accuracy, err, sizeof(err))) {
    set_last_error(err);
    libECP_free(h);
    return NULL;
  } else {
    h->st = &h->st_own;
  }
  g_phase[PH_TABLES] = now_ms() - t0;
  return h;
}

/* The device context is created on the first calculation, so that handles
   (and their host-side screening) exist without a GPU; the calculation
   itself fails loudly when no HIP device is present. */
static int ensure_device(libECPHandle *h) {
  char err[512] = "";
  if (h->dev) return 0;
  h->dev_id = ecp_dev_current();
  if (h->slot) h->dev = cache_take(h->slot, h->dev_id);
  if (!h->dev) h->dev = ecp_dev_create(h->st, err, sizeof(err));
  if (!h->dev) {
    set_last_error(err);
    return -1;
  }
  return 0;
}

/* Evaluates every triplet of the handle's molecule on the device and fills
   h->blocks / h->status; returns 0, or -1 on a device/host error. */
static int run_handle(libECPHandle *h) {
  char err[512] = "";
  for (int i = PH_SCREEN; i < PH_N; i++) g_phase[i] = 0.0;
  double t = now_ms();
  if (h->job_ready) {
    ecp_job_free(&h->job);
    h->job_ready = 0;
  }
  if (ecp_job_build(&h->mol, 1, h->n, h->st, &h->job, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  h->job_ready = 1;
  if (h->job.nblk > h->cap_blocks) {
    free(h->blocks);
    h->blocks = malloc(sizeof(double) * h->job.nblk);
    h->cap_blocks = h->blocks ? h->job.nblk : 0;
  }
  if (2 * (size_t)h->job.ntrip > h->cap_status) {
    free(h->status);
    h->status = malloc(2 * (size_t)h->job.ntrip);
    h->cap_status = h->status ? 2 * (size_t)h->job.ntrip : 0;
  }
  if ((h->job.nblk && !h->blocks) || (h->job.ntrip && !h->status)) {
    set_last_error("out of host memory for the blocks");
    return -1;
  }
  double t1 = now_ms();
  g_phase[PH_SCREEN] = t1 - t;
  if (h->job.ntrip == 0) return 0;
  if (ensure_device(h)) return -1;
  t = now_ms();
  g_phase[PH_CONTEXT] = t - t1;
  /* prepare (host work lists, layout, upload) / execute (kernels, waits) /
     fetch (download, waits); the wait after prepare only separates the
     upload from the kernels in the phase timings */
  if (ecp_dev_prepare(h->dev, &h->job, err, sizeof(err)) || ecp_dev_sync(h->dev)) {
    set_last_error(err[0] ? err : "device error after the upload");
    return -1;
  }
  t1 = now_ms();
  g_phase[PH_UPLOAD] = t1 - t;
  if (ecp_dev_execute(h->dev, 1, NULL, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  t = now_ms();
  g_phase[PH_KERNELS] = t - t1;
  if (ecp_dev_fetch(h->dev, h->blocks, h->status, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  g_phase[PH_DOWNLOAD] = now_ms() - t;
  return 0;
}

static int deliver(libECPHandle *h, ECPCallback cb, void *args);

int calculateECPIntegrals(libECPHandle *h, ECPCallback cb, void *args) {
  if (!h) {
    set_last_error("NULL handle");
    return -1;
  }
  if (run_handle(h)) return -1;
  const double t0 = now_ms();
  const int rc = deliver(h, cb, args);
  g_phase[PH_DELIVER] = now_ms() - t0;
  return rc;
}

static int deliver(libECPHandle *h, ECPCallback cb, void *args) {
  const ecp_job *J = &h->job;
  /* callbacks in the reference order (libecp.c:372); stop at the first
     failed integration with its documented code (libecp.h:23-26) */
  for (int t = 0; t < J->ntrip; t++) {
    const ecp_triplet *tr = &J->trip[t];
    const ecp_shell *Sa = &J->shells[tr->sa], *Sb = &J->shells[tr->sb];
    const int A = Sa->atom, B = Sb->atom, C = J->cent[tr->centre].atom;
    const int sa = tr->sa - J->atom_shell0[A], sb = tr->sb - J->atom_shell0[B];
    if (h->status[2 * t]) return 1;
    if (cb) cb(A, sa, Sa->l, tr->shifta, B, sb, Sb->l, tr->shiftb, C, h->blocks + tr->blk1, args);
    if (h->status[2 * t + 1]) return 2;
    if (cb) cb(A, sa, Sa->l, tr->shifta, B, sb, Sb->l, tr->shiftb, C, h->blocks + tr->blk2, args);
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* getIntegrals (getIntegrals.c:22-95): upper-triangle += of every block. */

int getIntegrals(int nrAtoms, double *geometry, int *shellsECP, int *KECP, int *lECP, double *nECP, double *dECP,
                 double *aECP, int *shellsBS, int *lBS, int *KBS, double *dBS, double *aBS, int largeGridOrder,
                 double tolerance, double accuracy, int rowdim, double *I) {
  libECPHandle *h = libECP_init(nrAtoms, geometry, shellsECP, lECP, KECP, nECP, dECP, aECP, shellsBS, lBS, KBS,
                                dBS, aBS, 0, -1, NULL, largeGridOrder, tolerance, accuracy);
  if (!h) {
    printf("error initializing libECP\n");
    return 1;
  }
  if (run_handle(h)) {
    fprintf(stderr, "libecp: %s\n", g_err);
    libECP_free(h);
    return 1;
  }
  const double t0 = now_ms();
  if (h->job.ntrip > 0) {
    int *aoff = malloc(sizeof(int) * ((size_t)h->job.nsh + 1));
    if (!aoff) {
      fprintf(stderr, "libecp: out of host memory\n");
      libECP_free(h);
      return 1;
    }
    ecp_job_ao_offsets(&h->job, aoff);
    /* the reference stops accumulating at a failed integration */
    (void)ecp_job_scatter(&h->job, 0, aoff, h->blocks, h->status, rowdim, 0, I);
    free(aoff);
  }
  g_phase[PH_DELIVER] = now_ms() - t0;
  libECP_free(h);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Density contraction on the device (include/libecp_ext.h).              */

int libecp_contract(libECPHandle *h, int rowdim, const double *P, double *out) {
  char err[512] = "";
  if (!h || !P || !out) {
    set_last_error("NULL argument");
    return -1;
  }
  if (h->n > 1) {
    set_last_error("density contraction supports derivative orders 0 and 1");
    return -1;
  }
  if (h->job_ready) {
    ecp_job_free(&h->job);
    h->job_ready = 0;
  }
  if (ecp_job_build(&h->mol, 1, h->n, h->st, &h->job, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  h->job_ready = 1;
  int nao = 0;
  for (int s = 0; s < h->job.nsh; s++) nao += ECP_DIM_IJK(h->job.shells[s].l);
  if (rowdim < nao) {
    set_last_error("rowdim is smaller than the number of basis functions");
    return -1;
  }
  const int nout = h->n == 0 ? 1 : 3 * h->nat;
  if (h->job.ntrip == 0) {
    for (int i = 0; i < nout; i++) out[i] = 0.0;
    return 0;
  }
  if (ensure_device(h)) return -1;
  const long long poff = 0;
  const int prow = rowdim;
  const size_t nP = nao ? (size_t)rowdim * (nao - 1) + nao : 0;
  int code = 0;
  if (ecp_dev_prepare(h->dev, &h->job, err, sizeof(err)) ||
      ecp_dev_execute(h->dev, 1, NULL, err, sizeof(err)) ||
      ecp_dev_contract(h->dev, &h->job, P, nP, 0, &poff, &prow, out, 0, &code, NULL, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  return code;
}

/* ---------------------------------------------------------------------- */
/* Additive queries (include/libecp_ext.h).                               */

int libecp_screened_triplets(libECPHandle *h, int *keys, int cap) {
  char err[512] = "";
  if (!h) return -1;
  if (h->job_ready) {
    ecp_job_free(&h->job);
    h->job_ready = 0;
  }
  if (ecp_job_build(&h->mol, 1, h->n, h->st, &h->job, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  h->job_ready = 1;
  const ecp_job *J = &h->job;
  for (int t = 0; t < J->ntrip && keys && t < cap; t++) {
    const ecp_triplet *tr = &J->trip[t];
    const int A = J->shells[tr->sa].atom, B = J->shells[tr->sb].atom;
    keys[5 * t + 0] = J->cent[tr->centre].atom;
    keys[5 * t + 1] = A;
    keys[5 * t + 2] = tr->sa - J->atom_shell0[A];
    keys[5 * t + 3] = B;
    keys[5 * t + 4] = tr->sb - J->atom_shell0[B];
  }
  return J->ntrip;
}

int libecp_model_flops(libECPHandle *h, double *out, int n) {
  char err[512] = "";
  if (!h || !out) {
    set_last_error("NULL argument");
    return -1;
  }
  if (h->job_ready) {
    ecp_job_free(&h->job);
    h->job_ready = 0;
  }
  if (ecp_job_build(&h->mol, 1, h->n, h->st, &h->job, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  h->job_ready = 1;
  double v[ECP_FL_N];
  ecp_model_flops(&h->job, v);
  for (int i = 0; i < ECP_FL_N && i < n; i++) out[i] = v[i];
  return ECP_FL_N;
}

double libecp_model_fallback_flops(int type, int la, int lb, int l, int ng_l, int failed, int points) {
  return type == 1 ? ecp_model_fb1_unit(la + lb, failed, points) : ecp_model_fb2_pair(la, lb, l, ng_l, failed, points);
}

long long libecp_nominal_triplets(libECPHandle *h) {
  if (!h) return -1;
  long long ne = 0, ns = 0;
  for (int i = 0; i < h->nat; i++) {
    ne += h->shellsECP[i] > 0;
    ns += h->shellsBS[i];
  }
  return ne * ns * (ns + 1) / 2;
}

/* ---------------------------------------------------------------------- */
/* Batched evaluation (include/libecp_ext.h).                              */

struct libecp_batch {
  int nmol, n;
  int *nao, *nat;  /* per molecule */
  ecp_mol *mols; /* private copies */
  ecp_static st;
  ecp_dev *dev;
  ecp_job job;
  int *aoff;
  double *blocks;
  unsigned char *status;
  long long nominal;
  char *con_img;        /* contraction index tables of the packed density layout */
  size_t con_img_bytes;
};

void libecp_batch_free(libecp_batch *b) {
  if (!b) return;
  ecp_dev_destroy(b->dev);
  ecp_job_free(&b->job);
  ecp_static_free(&b->st);
  for (int i = 0; i < b->nmol; i++) ecp_mol_release(&b->mols[i]);
  free(b->mols);
  free(b->aoff);
  free(b->nao);
  free(b->nat);
  free(b->blocks);
  free(b->status);
  free(b->con_img);
  free(b);
}

libecp_batch *libecp_batch_create(const libecp_molecule *mols, int nmol, int largeGridOrder, double tolerance,
                                  double accuracy) {
  return libecp_batch_create_n(mols, nmol, 0, largeGridOrder, tolerance, accuracy);
}

libecp_batch *libecp_batch_create_n(const libecp_molecule *mols, int nmol, int n, int largeGridOrder,
                                    double tolerance, double accuracy) {
  char err[512] = "";
  set_last_error("");
  if (nmol <= 0 || !mols) {
    set_last_error("empty batch");
    return NULL;
  }
  if (n < 0 || n > 2) {
    set_last_error("derivative order must be 0, 1 or 2");
    return NULL;
  }
  libecp_batch *b = calloc(1, sizeof(*b));
  if (!b) {
    set_last_error("out of host memory");
    return NULL;
  }
  b->n = n;
  b->nao = calloc((size_t)nmol, sizeof(int));
  b->nat = calloc((size_t)nmol, sizeof(int));
  b->mols = calloc((size_t)nmol, sizeof(ecp_mol));
  if (!b->nao || !b->nat || !b->mols) {
    set_last_error("out of host memory");
    libecp_batch_free(b);
    return NULL;
  }
  b->nmol = nmol;
  int maxL = 1, maxLBS = 0;
  for (int i = 0; i < nmol; i++) {
    const libecp_molecule *M = &mols[i];
    ecp_mol m0 = {M->nrAtoms, M->geometry, M->shellsECP, M->lECP, M->KECP, M->nECP, M->dECP, M->aECP,
                  M->shellsBS, M->lBS, M->KBS, M->dBS, M->aBS};
    ecp_mol_dims d;
    if (ecp_mol_copy(&m0, &b->mols[i], &d, err, sizeof(err))) {
      char msg[600];
      snprintf(msg, sizeof(msg), "molecule %d: %s", i, err);
      set_last_error(msg);
      libecp_batch_free(b);
      return NULL;
    }
    b->nominal += (long long)d.necp * d.nsh * (d.nsh + 1) / 2;
    b->nao[i] = d.nao;
    b->nat[i] = M->nrAtoms;
    if (d.maxLECP > maxL) maxL = d.maxLECP;
    if (d.maxLBS > maxLBS) maxLBS = d.maxLBS;
  }
  if (ecp_static_init(&b->st, maxL, maxLBS + n, NULL, -1, largeGridOrder, tolerance, accuracy, err, sizeof(err)) ||
      ecp_job_build(b->mols, nmol, n, &b->st, &b->job, err, sizeof(err))) {
    set_last_error(err);
    libecp_batch_free(b);
    return NULL;
  }
  b->aoff = malloc(sizeof(int) * ((size_t)b->job.nsh + 1));
  b->blocks = malloc(sizeof(double) * (b->job.nblk + 1));
  b->status = calloc(2 * (size_t)b->job.ntrip + 1, 1);
  if (!b->aoff || !b->blocks || !b->status) {
    set_last_error("out of host memory for the batch's blocks");
    libecp_batch_free(b);
    return NULL;
  }
  ecp_job_ao_offsets(&b->job, b->aoff);
  if (n <= 1) {
    /* the index tables of libecp_batch_contract_noalloc (densities packed
       by molecule), built here so that the call itself allocates nothing */
    long long *poff = malloc(sizeof(long long) * (size_t)nmol);
    int *prow = malloc(sizeof(int) * (size_t)nmol);
    size_t nP = 0;
    if (poff && prow) {
      for (int m = 0; m < nmol; m++) {
        poff[m] = (long long)nP;
        prow[m] = b->nao[m];
        nP += (size_t)b->nao[m] * b->nao[m];
      }
      b->con_img_bytes = ecp_dev_contract_image(&b->job, poff, prow, NULL, 0);
      b->con_img = malloc(b->con_img_bytes ? b->con_img_bytes : 1);
      if (b->con_img) ecp_dev_contract_image(&b->job, poff, prow, b->con_img, b->con_img_bytes);
    }
    free(poff);
    free(prow);
    if (!b->con_img) {
      set_last_error("out of host memory for the batch's contraction tables");
      libecp_batch_free(b);
      return NULL;
    }
  }
  b->dev = ecp_dev_create(&b->st, err, sizeof(err));
  if (!b->dev || ecp_dev_prepare(b->dev, &b->job, err, sizeof(err))) {
    set_last_error(err);
    libecp_batch_free(b);
    return NULL;
  }
  return b;
}

int libecp_batch_execute(libecp_batch *b, int repeats, double *ms) {
  char err[512] = "";
  if (!b) return -1;
  if (ecp_dev_execute(b->dev, repeats, ms, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  return 0;
}

int libecp_batch_fetch(libecp_batch *b) {
  char err[512] = "";
  if (!b) return -1;
  if (ecp_dev_fetch(b->dev, b->blocks, b->status, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  return 0;
}

int libecp_batch_matrix(libecp_batch *b, int mol, int rowdim, double *I) {
  if (!b || mol < 0 || mol >= b->nmol) return -1;
  if (b->n != 0) {
    set_last_error("the ECP matrix of a derivative batch is not defined (use libecp_batch_contract)");
    return -1;
  }
  return ecp_job_scatter(&b->job, mol, b->aoff, b->blocks, b->status, rowdim, 0, I);
}

int libecp_batch_contract_noalloc(libecp_batch *b, const double *P, double *out, int *codes, double *ms) {
  char err[512] = "";
  if (!b || !P || !out) {
    set_last_error("NULL argument");
    return -1;
  }
  if (b->n > 1 || !b->con_img) {
    set_last_error("density contraction supports derivative orders 0 and 1");
    return -1;
  }
  if (ecp_dev_contract_noalloc(b->dev, &b->job, b->con_img, b->con_img_bytes, P, out, codes, ms, err,
                               sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  return 0;
}

int libecp_batch_contract(libecp_batch *b, const double *P, double *out, int *codes, int device, double *ms) {
  char err[512] = "";
  if (!b || !P || !out) {
    set_last_error("NULL argument");
    return -1;
  }
  if (b->n > 1) {
    set_last_error("density contraction supports derivative orders 0 and 1");
    return -1;
  }
  long long *poff = malloc(sizeof(long long) * (size_t)b->nmol);
  int *prow = malloc(sizeof(int) * (size_t)b->nmol);
  int *cd = codes ? codes : malloc(sizeof(int) * (size_t)b->nmol);
  if (!poff || !prow || !cd) {
    set_last_error("out of host memory");
    free(poff);
    free(prow);
    if (cd != codes) free(cd);
    return -1;
  }
  size_t nP = 0;
  for (int m = 0; m < b->nmol; m++) {
    poff[m] = (long long)nP;
    prow[m] = b->nao[m];
    nP += (size_t)b->nao[m] * b->nao[m];
  }
  int rc = 0;
  if (ecp_dev_contract(b->dev, &b->job, P, nP, device, poff, prow, out, device, cd, ms, err, sizeof(err))) {
    set_last_error(err);
    rc = -1;
  }
  free(poff);
  free(prow);
  if (!codes) free(cd);
  return rc;
}

size_t libecp_batch_block_doubles(libecp_batch *b) { return b ? b->job.nblk : 0; }

long long libecp_batch_triplets(libecp_batch *b, long long *evaluated) {
  if (!b) return -1;
  if (evaluated) *evaluated = b->job.ntrip;
  return b->nominal;
}

int libecp_batch_stage_ms(libecp_batch *b, double *ms, int n) {
  if (!b || !b->dev) return 0;
  return ecp_dev_last_timing(b->dev, ms, n);
}

int libecp_batch_quadrature_ms(libecp_batch *b, double *ms, int *samples) {
  char err[512] = "";
  if (!b || !ms) return -1;
  if (ecp_dev_span_ms(b->dev, ms, samples, err, sizeof(err))) {
    set_last_error(err);
    return -1;
  }
  return 0;
}

int libecp_batch_profile(libecp_batch *b, int *launches, double *ms, double *flops, int n) {
  char err[512] = "";
  int l[16];
  double m[16], f[16];
  if (!b) return -1;
  int k = ecp_dev_profile(b->dev, l, m, f, err, sizeof(err));
  if (k < 0) {
    set_last_error(err);
    return -1;
  }
  for (int i = 0; i < k && i < n; i++) {
    launches[i] = l[i];
    ms[i] = m[i];
    flops[i] = f[i];
  }
  return k;
}

const char *libecp_kernel_family(int i) { return ecp_dev_family_name(i); }

/* ---------------------------------------------------------------------- */
/* Launch parameters and autotuning of a batch (include/libecp_ext.h).     */

static int *tuning_field(ecp_tuning *t, const char *name) {
  if (!strcmp(name, "t1_head_coop")) return &t->t1_head_coop;
  if (!strcmp(name, "t1_head_serial")) return &t->t1_head_serial;
  if (!strcmp(name, "t1_coop_max")) return &t->t1_coop_max;
  if (!strcmp(name, "t1_tail_grid")) return &t->t1_tail_grid;
  if (!strcmp(name, "fuse")) return &t->fuse;
  if (!strcmp(name, "asm2_cache")) return &t->asm2_cache;
  if (!strcmp(name, "t2_head")) return &t->t2_head;
  if (!strcmp(name, "t2_tail_grid")) return &t->t2_tail_grid;
  if (!strcmp(name, "fb2_split")) return &t->fb2_split;
  if (!strcmp(name, "t1_spec")) return &t->t1_spec;
  if (!strcmp(name, "t2_spec")) return &t->t2_spec;
  return NULL;
}

int libecp_batch_set_param(libecp_batch *b, const char *name, int value) {
  ecp_tuning t;
  if (!b || !name) return -1;
  ecp_dev_get_tuning(b->dev, &t);
  int *f = tuning_field(&t, name);
  if (!f) {
    set_last_error("unknown launch parameter");
    return -1;
  }
  *f = value;
  ecp_dev_set_tuning(b->dev, &t);
  return 0;
}

int libecp_batch_get_param(libecp_batch *b, const char *name, int *value) {
  ecp_tuning t;
  if (!b || !name || !value) return -1;
  ecp_dev_get_tuning(b->dev, &t);
  int *f = tuning_field(&t, name);
  if (!f) {
    set_last_error("unknown launch parameter");
    return -1;
  }
  *value = *f;
  return 0;
}

/* device time (ms) of one evaluation under tuning t: graph replays */
static double time_tuning(libecp_batch *b, const ecp_tuning *t, int repeats, char *err, size_t errlen) {
  double ms = 0.0;
  ecp_dev_set_tuning(b->dev, t);
  if (ecp_dev_execute(b->dev, 1 + repeats, &ms, err, errlen)) return -1.0;
  return ms;
}

/* Coordinate search over the launch parameters, one at a time, keeping a
   value only when it is faster than the best so far by more than 1%. */
int libecp_batch_autotune(libecp_batch *b, int repeats, char *report, int len) {
  char err[512] = "";
  if (!b) return -1;
  if (repeats < 1) repeats = 3;
  ecp_tuning best;
  ecp_dev_get_tuning(b->dev, &best);
  double tbest = time_tuning(b, &best, repeats, err, sizeof(err));
  if (tbest < 0) {
    set_last_error(err);
    return -1;
  }
  size_t used = 0;
  if (report && len > 0) {
    report[0] = 0;
    used = (size_t)snprintf(report, (size_t)len, "start %.4f ms", tbest);
  }
  static const struct {
    const char *name;
    int n;
    int v[6];
  } space[] = {
      {"fuse", 3, {0, 1, 2}},
      {"asm2_cache", 3, {0, 1, 2}},
      {"t2_head", 6, {0, 4, 6, 8, 10, 15}},
      {"t1_coop_max", 5, {0, 20000, 100000, 400000, 1 << 30}},
      {"t1_head_coop", 4, {6, 8, 10, 12}},
      {"t1_head_serial", 4, {6, 8, 10, 15}},
      {"t1_tail_grid", 4, {2048, 8192, 32768, 131072}},
      {"fb2_split", 4, {0, 1, 2, 3}},
      {"t1_spec", 2, {0, 1}},
  };
  for (size_t k = 0; k < sizeof(space) / sizeof(space[0]); k++) {
    for (int i = 0; i < space[k].n; i++) {
      ecp_tuning t = best;
      int *f = tuning_field(&t, space[k].name);
      if (*f == space[k].v[i]) continue;
      *f = space[k].v[i];
      const double ms = time_tuning(b, &t, repeats, err, sizeof(err));
      if (ms < 0) {
        set_last_error(err);
        ecp_dev_set_tuning(b->dev, &best);
        return -1;
      }
      if (report && used < (size_t)len)
        used += (size_t)snprintf(report + used, (size_t)len - used, "; %s=%d %.4f ms", space[k].name,
                                 space[k].v[i], ms);
      if (ms < 0.99 * tbest) {
        tbest = ms;
        best = t;
      }
    }
  }
  ecp_dev_set_tuning(b->dev, &best);
  if (report && used < (size_t)len)
    snprintf(report + used, (size_t)len - used, "; best %.4f ms: fuse=%d asm2_cache=%d t2_head=%d t1_coop_max=%d "
             "t1_head_coop=%d t1_head_serial=%d t1_tail_grid=%d", tbest, best.fuse, best.asm2_cache, best.t2_head,
             best.t1_coop_max, best.t1_head_coop, best.t1_head_serial, best.t1_tail_grid);
  return 0;
}
