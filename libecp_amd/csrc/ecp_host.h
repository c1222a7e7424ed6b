/* Internal interface between the C host library and the HIP device layer. */
#ifndef ECP_HOST_H
#define ECP_HOST_H 1

#include <stddef.h>
#include "ecp_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables (ecp_tables.c) ---- */
double *ecp_factorials(int n, int step);
int *ecp_cart_order(int am);
int *ecp_cart_index(int am, const int *ijk);
int ecp_tm_offset(int l);
double *ecp_cart2sph(int lmax, const int *xyz, const double *fac);
double *ecp_poly2sph(const double *M, int lmax, const int *xyz, const double *fac2);
int ecp_grid_psm92(int maxPoints, double *x, double *w);
int ecp_grid_ps93_kk(double *x, double *w);
void ecp_schedule_ps93(ecp_schedule *S, int *pairs_l, int *pairs_r);
void ecp_schedule_psm92(ecp_schedule *S, int order, int *pairs_l, int *pairs_r);
int ecp_bessel_table(double accuracy, int N, int cutoff, int rows, double *Kdev);
double *ecp_omega(int maxL, int maxLambda, int maxAlpha, const int *ijk, const int *ijkIndex, int ijkDim,
                  const double *cart2sph, const double *poly2sph, int tmDim, const double *dfac);

/* ---- static (geometry-independent) tables shared by every job ---- */
typedef struct {
  double xs[ECP_SMALL_ORDER], ws[ECP_SMALL_ORDER];
  double xl[ECP_LARGE_MAX], wl[ECP_LARGE_MAX];
  int lorder;
  ecp_schedule s93, s92;
  int p93l[ECP_SMALL_ORDER], p93r[ECP_SMALL_ORDER];
  int p92l[ECP_LARGE_MAX], p92r[ECP_LARGE_MAX];
  double *bessel;            /* (ECP_BESSEL_N+1) x ECP_BESSEL_STRIDE */
  double *fac, *dfac;
  int *ijk, *ijkIndex, ijkDim;
  int tmDim;
  double *cart2sph, *poly2sph;
  int omL, omLambda, omAlpha; /* dims of the Omega table */
  double *omega;
  double tolerance, accuracy, lnAcc1, lnAcc2;
} ecp_static;

/* ---- per-call job (flat arrays, host side) ---- */
typedef struct {
  int nmol;
  int deriv;                 /* derivative order n (0, 1, 2) */
  int nat, nsh, nprim, ncent, ngauss;
  ecp_xyz *atoms;            /* [nat] */
  int *atom_maxl;            /* [nat] max basis l on the atom */
  int *atom_shell0, *atom_nsh; /* [nat] first global shell, count */
  ecp_shell *shells;         /* [nsh] */
  double *pa, *pd;           /* [nprim] exponents, coefficients */
  ecp_centre *cent;          /* [ncent] */
  ecp_gauss *gauss;          /* [ngauss] */
  ecp_window *win;           /* per centre, per shell of its molecule */
  int nwin;
  int ntrip;
  ecp_triplet *trip;         /* [ntrip], per molecule in the reference order */
  int *trip_mol_begin;       /* [nmol+1] */
  int *mol_atom0;            /* [nmol+1] first global atom of each molecule */
  size_t nblk;               /* doubles of output blocks */
  int maxK;                  /* largest contraction length */
  int maxL;                  /* largest ECP local label */
  int FL;                    /* F rows per (centre, shell, run) slab; runs 0..deriv */
  int UK;                    /* r-power rows of the Ut table */
  size_t nUL, nUt, nF, nang; /* table sizes (doubles) */
  int angL;                  /* l-dimension of the per-(centre, atom) angular table */
  int ang_lam, ang_c;        /* its lambda and Cartesian dimensions */
} ecp_job;

/* ---- launch parameters of the device pipeline (tunable per context) ---- */
typedef struct {
  int t1_head_coop;   /* PS93 levels of the type-1 head kernel before the wave-cooperative tail */
  int t1_head_serial; /* ... before the lane-serial tail (>= the level count: no tail) */
  int t1_coop_max;    /* largest class (primitive pairs) that takes the cooperative tail */
  int t1_tail_grid;   /* most workgroups of a tail launch (persistent over its queue) */
  int fuse;           /* 1: one launch per stage over all classes, assembly overlapping the
                         large-grid fallbacks; 0: per-class launches on forked streams;
                         2: fused for small jobs, per-class for large ones */
  int asm2_cache;     /* type-2 assembly with angular factors in registers: 0 no, 1 yes
                         (the default), 2 with fused launches only */
  int t2_head;        /* PS93 levels of the type-2 head kernel before the compacted tail
                         (>= the level count, the default: no tail; 0: 6 with fused
                         launches, else none) */
  int t2_tail_grid;   /* most workgroups of the type-2 tail launch */
  int fb2_split;      /* waves per (triplet, primitive pair) of the type-2 fallback, its 64-entry
                         passes dealt round robin; 0: 3 for small jobs, 1 for large ones */
  int t1_spec;        /* fused launches: the type-1 large-grid fallback starts after the head,
                         beside the tail, on every entry the head left open (1), or after the
                         tail on the failed ones only (0) */
  int t2_spec;        /* fused launches with a type-2 head: the type-2 fallback starts after
                         the head, beside the tail, on every entry the head left open (1), or
                         after the tail on the failed ones only (0) */
} ecp_tuning;

/* ---- device layer (ecp_device.hip) ---- */
typedef struct ecp_dev ecp_dev;

ecp_dev *ecp_dev_create(const ecp_static *st, char *err, size_t errlen);
void ecp_dev_destroy(ecp_dev *d);
/* Runs the whole pipeline for a job: returns 0 and fills blocks (job->nblk
   doubles) and status (2 bytes per triplet: type-1 / type-2 failure flags),
   or a negative value on a device error (message in err). */
int ecp_dev_run(ecp_dev *d, const ecp_job *job, double *blocks, unsigned char *status, char *err,
                size_t errlen);
/* The same in three steps: upload a job (inputs become resident in HBM),
   execute its kernel sequence `repeats` times (graph replays after the
   first; *ms_total = device time per repeat), and copy the results back. */
int ecp_dev_prepare(ecp_dev *d, const ecp_job *job, char *err, size_t errlen);
/* prepare = plan + bind.  plan does the host work (work lists, layout,
   staging in pinned memory) and returns the bytes of device memory the job
   needs; bind places the job at mem (capacity cap; NULL = the context's own
   arena) and uploads its inputs. */
int ecp_dev_plan(ecp_dev *d, const ecp_job *job, size_t *bytes, char *err, size_t errlen);
int ecp_dev_bind(ecp_dev *d, void *mem, size_t cap, char *err, size_t errlen);
int ecp_dev_execute(ecp_dev *d, int repeats, double *ms_total, char *err, size_t errlen);
int ecp_dev_fetch(ecp_dev *d, double *blocks, unsigned char *status, char *err, size_t errlen);
/* Contracts the last evaluation's blocks with densities on the device (job
   of derivative order 0: one energy per molecule; order 1: 3 gradient
   components per atom, molecule after molecule).  P holds nP doubles,
   molecule m's symmetric density at poff[m] with row stride prow[m]; P and
   out are device pointers when the *_on_device flags are set.  codes[m] =
   0, or 1 / 2 when a type-1 / type-2 integration of molecule m failed.
   *ms = device time of the contraction kernels.  0, or -1 (message in err). */
int ecp_dev_contract(ecp_dev *d, const ecp_job *job, const double *P, size_t nP, int P_on_device,
                     const long long *poff, const int *prow, double *out, int out_on_device, int *codes,
                     double *ms, char *err, size_t errlen);
/* The contraction's index tables for densities at poff / prow (the layout
   ecp_dev_contract uploads): returns their size in bytes and writes them to
   img when cap is large enough. */
size_t ecp_dev_contract_image(const ecp_job *job, const long long *poff, const int *prow, char *img, size_t cap);
/* ecp_dev_contract with device P and out and no allocation: the index
   tables (img, from ecp_dev_contract_image), partial sums and codes go to
   the device's static scratch (32 MiB; larger jobs are rejected with a
   message).  Calls that overlap on one device share the scratch and race.
   codes (host, may be NULL).  0, or -1 (message in err). */
int ecp_dev_contract_noalloc(ecp_dev *d, const ecp_job *job, const char *img, size_t img_bytes, const double *P,
                             double *out, int *codes, double *ms, char *err, size_t errlen);
/* One eager evaluation timed per kernel launch (HIP events on the launch's
   own stream).  Per kernel family (tables, type1, type2, fallback1,
   fallback2, assemble): launches, summed duration in ms, algorithmic FLOPs
   of one evaluation (type1/type2 only).  Returns the number of families. */
int ecp_dev_profile(ecp_dev *d, int *launches, double *ms, double *flops, char *err, size_t errlen);
const char *ecp_dev_family_name(int f);
/* Stage timings of the last execute (ms): tables, quadratures (small grid
   and large-grid fallbacks, concurrent), assembly[, mean graph replay]. */
int ecp_dev_last_timing(ecp_dev *d, double *ms, int n);
/* Device time (ms) from the start of the type-1 head launch to the end of
   the type-1 tail launch (events on their stream, recorded inside graph
   replays too): the last execute's eager evaluation, or the mean over its
   last replays (up to 8, one per captured graph copy; *samples = count);
   fused launches only, else -1. */
int ecp_dev_span_ms(ecp_dev *d, double *ms, int *samples, char *err, size_t errlen);
/* launch parameters: defaults at creation; setting them drops a captured
   graph (the next execute runs eagerly and captures again) */
void ecp_dev_get_tuning(const ecp_dev *d, ecp_tuning *t);
void ecp_dev_set_tuning(ecp_dev *d, const ecp_tuning *t);
/* Waits for every stream of the context (after a failed bind / execute /
   fetch, before its device memory is reused).  0, or -1 when a stream
   reports an error. */
int ecp_dev_sync(ecp_dev *d);
/* device memory and the calling thread's current HIP device */
int ecp_dev_mem_alloc(void **p, size_t bytes, char *err, size_t errlen);
void ecp_dev_mem_free(void *p);
int ecp_dev_current(void);
int ecp_dev_use(int device);

/* ---- molecules and jobs (ecp_job.c) ---- */
typedef struct {
  int nat;
  const double *geom;
  const int *shellsECP, *lECP, *KECP;
  const double *nECP, *dECP, *aECP;
  const int *shellsBS, *lBS, *KBS;
  const double *dBS, *aBS;
} ecp_mol;

typedef struct {
  int maxLECP, maxLBS, nsh, nprim, ngauss, necp, nao;
} ecp_mol_dims;

int ecp_mol_scan(const ecp_mol *m, ecp_mol_dims *d, char *err, size_t errlen);
int ecp_static_init(ecp_static *st, int maxL, int maxAlpha, const int *ordering, int lmaxOrdering,
                    int largeGridOrder, double tolerance, double accuracy, char *err, size_t errlen);
void ecp_static_free(ecp_static *st);
int ecp_job_build(const ecp_mol *mols, int nmol, int deriv, const ecp_static *st, ecp_job *job, char *err,
                  size_t errlen);
void ecp_job_free(ecp_job *job);
int ecp_mol_copy(const ecp_mol *src, ecp_mol *dst, ecp_mol_dims *d, char *err, size_t errlen);
void ecp_mol_release(ecp_mol *m);
void ecp_job_ao_offsets(const ecp_job *J, int *aoff);
int ecp_job_scatter(const ecp_job *J, int m, const int *aoff, const double *blocks, const unsigned char *status,
                    int rowdim, int packed_n, double *I);

/* ---- algorithmic FLOP model (ecp_flops.c) ---- */
enum { ECP_FL_FTAB = 0, ECP_FL_T1, ECP_FL_T2, ECP_FL_ASM1, ECP_FL_ASM2, ECP_FL_SHIFT, ECP_FL_FB1, ECP_FL_FB2,
       ECP_FL_N };
/* the terms known before evaluation (all but the fallbacks) of a job */
void ecp_model_flops(const ecp_job *J, double *out /* [ECP_FL_N] */);
double ecp_model_t1_unit(int lab, int G);
double ecp_model_asm1_unit(int la, int lb);
double ecp_model_t2_group(int la, int lb, int l, int G);
double ecp_model_asm2_group(int la, int lb, int l);
double ecp_model_shift(int la, int lb);
double ecp_model_ftab(int K, int G, int m);
/* fallbacks: a primitive pair of class lab with f failed entries, N points;
   a (triplet, l) group with f failed entries, per primitive pair */
double ecp_model_fb1_unit(int lab, int f, int N);
double ecp_model_fb2_pair(int la, int lb, int l, int ng_l, int f, int N);
int ecp_model_gauss_of(const ecp_job *J, int centre, int l);

/* ---- host helpers (ecp_job.c) ---- */
double ecp_eval_ecp_host(const ecp_gauss *g, int ng, int l, double r);
double ecp_shell_radius(int K, int l, const double *d, const double *a, double zero, int *ok);

#ifdef __cplusplus
}
#endif
#endif
