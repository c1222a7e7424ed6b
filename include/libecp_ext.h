/* libecp_ext.h -- additive entry points of the MI355X ECP-integral library.
 *
 * None of these has a counterpart in the reference (chrr/libECP); they sit
 * beside the drop-in API of libecp.h / getIntegrals.h and never change it.
 */
#ifndef LIBECP_EXT_H
#define LIBECP_EXT_H 1

#include <stddef.h>

#include "libecp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host-side screening of a handle's molecule (no device work): returns the
   number of (C, shell a, shell b >= a) triplets that calculateECPIntegrals
   evaluates, in its callback order; when keys != NULL, writes up to cap
   records of 5 ints (C, A, sa, B, sb).  Returns -1 on error. */
int libecp_screened_triplets(libECPHandle *h, int *keys, int cap);

/* N_ECP * N_sh (N_sh + 1) / 2: the reference's triplet loop count before
   screening (libecp.c:256-312), the unit of the throughput metric. */
long long libecp_nominal_triplets(libECPHandle *h);

/* Algorithmic FLOPs of the reference's eager algorithm on the handle's
   screened triplets (host only, no device work), up to n of: F tables,
   type-1 small grid, type-2 small grid, type-1 assembly, type-2 assembly,
   binomial shift (the fallbacks depend on which integrals fail, known only
   after an evaluation: libecp_batch_profile reports them).  FMA = 2,
   add/sub/mul/div = 1, exp/pow/sqrt/log/trig = 1.  Returns the number of
   terms (8; the last two, the fallbacks, are 0 here), or -1. */
int libecp_model_flops(libECPHandle *h, double *out, int n);
/* The fallback terms of the same model: type 1, a primitive pair of shells
   la, lb with `failed` failed small-grid entries on a large grid of
   `points` points; type 2, one primitive pair of a (triplet, channel l)
   whose channel has ng_l Gaussians, with `failed` failed T entries. */
double libecp_model_fallback_flops(int type, int la, int lb, int l, int ng_l, int failed, int points);

/* Phases of the last getIntegrals or calculateECPIntegrals call made by the
   calling thread, in ms, up to n of: host tables (libECP_init; ~0 when they
   come from the process cache), screening (triplet list), device context
   (created, or taken from the cache), plan + upload, kernels, download,
   matrix accumulation or callbacks.  Returns the number of phases (7). */
int libecp_last_call_ms(double *ms, int n);

/* Process-wide cache.  Handles with the default Cartesian ordering share
   the geometry-independent tables of their key (largest ECP label, largest
   basis l + n, grid order, tolerance, accuracy; at most 8 keys) and reuse
   device contexts (streams, pinned staging, tables resident in HBM; at most
   2 idle per key and device): libECP_free returns the context to the cache
   instead of destroying it, and the next handle of the key takes it, so
   repeated getIntegrals calls pay for neither.  Idle contexts keep their
   device and pinned memory until libecp_cache_clear() (which frees them) or
   the end of the process.  libecp_cache_enable(0) clears the cache and
   disables it (every handle then builds and frees its own); returns the
   previous setting. */
void libecp_cache_clear(void);
int libecp_cache_enable(int on);

/* ---- density contraction on the device ----------------------------------
   The consumer sketched in getIntegrals.c:15-18 (a density matrix and a
   gradient in place of the AO matrix), evaluated where the blocks are: no
   block leaves the device and no callback runs.  P is the full symmetric
   density (nao x nao, row stride rowdim) in getIntegrals' AO order (atoms,
   their shells, the shells' Cartesian components).  A handle of order
   n = 0 gives out[0] = E = sum_{mu nu} P_{mu nu} V_{mu nu}; n = 1 gives
   out[3 A + d] = dE/dR_{A,d} for every atom A (basis functions and ECP
   centre move with their atom).  Returns 0, 1 / 2 when a type-1 / type-2
   integration failed (out is then not meaningful), or -1
   (libecp_last_error(); also for n = 2). */
int libecp_contract(libECPHandle *h, int rowdim, const double *P, double *out);

/* ---- batched evaluation ------------------------------------------------
   Many molecules in one resident job: one upload, one kernel sequence over
   the triplets of every molecule (replayed from a captured HIP graph), one
   download.  Molecule arrays are copied at creation. */

typedef struct {
  int nrAtoms;
  const double *geometry;
  const int *shellsECP, *KECP, *lECP;   /* getIntegrals order: KECP before lECP */
  const double *nECP, *dECP, *aECP;
  const int *shellsBS, *lBS, *KBS;
  const double *dBS, *aBS;
} libecp_molecule;

typedef struct libecp_batch libecp_batch;

/* Screens every molecule on the host and uploads the job.  NULL on error
   (libecp_last_error()). */
libecp_batch *libecp_batch_create(const libecp_molecule *mols, int nmol, int largeGridOrder, double tolerance,
                                  double accuracy);
/* The same for derivative order n (0, 1 or 2; libecp.c:246-250): the
   blocks of the shifted shells, as a handle of order n delivers them.  A
   batch of order n > 0 has no ECP matrix (libecp_batch_matrix fails); its
   use is libecp_batch_contract (n = 1). */
libecp_batch *libecp_batch_create_n(const libecp_molecule *mols, int nmol, int n, int largeGridOrder,
                                    double tolerance, double accuracy);
/* Evaluates the whole batch `repeats` times on the device and waits;
   *ms = device time of one evaluation.  0 on success, -1 on a device error. */
int libecp_batch_execute(libecp_batch *b, int repeats, double *ms);
/* Copies the results of the last evaluation to the host. */
int libecp_batch_fetch(libecp_batch *b);
/* Adds molecule `mol`'s ECP matrix into the upper triangle of I (row stride
   rowdim), as getIntegrals does; returns 0, or 1 / 2 when a type-1 / type-2
   integration of that molecule failed (blocks up to the failure added). */
int libecp_batch_matrix(libecp_batch *b, int mol, int rowdim, double *I);
/* Contracts the last evaluation with one density per molecule on the
   device (see libecp_contract): molecule m's full symmetric nao_m x nao_m
   density, row-major, at P + sum_{i<m} nao_i^2.  out: n = 0: one energy per
   molecule; n = 1: 3 nrAtoms_m gradient components per molecule, molecules
   in order.  codes[m] (codes may be NULL): 0, or 1 / 2 when a type-1 /
   type-2 integration of molecule m failed (its out entries are then not
   meaningful).  device != 0: P and out are device pointers on the batch's
   device (no host copies).  *ms (ms may be NULL): device time of the
   contraction kernels.  0, or -1. */
int libecp_batch_contract(libecp_batch *b, const double *P, double *out, int *codes, int device, double *ms);
/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  libecp_batch_contract with device buffers (P, out on
   the batch's device; codes on the host, may be NULL) that calls no
   hipMalloc, hipFree or other allocating routine: the index tables are
   built when the batch is created, and the partial sums, codes and index
   tables live in a static device array of 32 MiB per GPU, held until the
   process ends.  Jobs needing more (about 4 million triplets for the
   energy, 1.3 million shifted triplets for the gradient) are rejected with
   a message.  Every call on a GPU shares that GPU's array: calls that
   overlap on one GPU (other threads, other batches) race and must be
   serialised by the caller.  Derivative orders 0 and 1.  Results equal
   libecp_batch_contract's bit for bit.  0, or -1 (libecp_last_error()). */
int libecp_batch_contract_noalloc(libecp_batch *b, const double *P, double *out, int *codes, double *ms);
/* Nominal (pre-screening) and evaluated triplet counts of the batch. */
long long libecp_batch_triplets(libecp_batch *b, long long *evaluated);
/* Doubles of the evaluated blocks (type 1 and type 2) the batch holds on
   the device. */
size_t libecp_batch_block_doubles(libecp_batch *b);
/* Stage timings (ms) of the first evaluation: tables, quadratures (small
   grid and large-grid fallbacks), assembly; returns the count written. */
int libecp_batch_stage_ms(libecp_batch *b, double *ms, int n);
/* One evaluation timed kernel by kernel (HIP events on each launch's
   stream).  For up to n kernel families (names: libecp_kernel_family(i)):
   number of launches, summed duration (ms), algorithmic FLOPs of one
   evaluation under the reference algorithm's per-unit models (0 where no
   model is defined).  Returns the number of families, or -1. */
int libecp_batch_profile(libecp_batch *b, int *launches, double *ms, double *flops, int n);
const char *libecp_kernel_family(int i);
/* Device time (ms) of the type-1 small-grid quadrature -- the head and tail
   launches, from an event pair on their stream -- in the last
   libecp_batch_execute call: its eager evaluation, or the mean over its
   last graph replays (up to 8; the graph is captured in copies with their
   own event pair and replays cycle over them).  *samples (may be NULL):
   evaluations averaged.  0, or -1 when the evaluation used per-class
   launches (the quadrature then spans several streams). */
int libecp_batch_quadrature_ms(libecp_batch *b, double *ms, int *samples);
/* Launch parameters of the batch's device pipeline, by name:
   "t1_head_coop" / "t1_head_serial" (PS93 levels the type-1 head kernel runs
   before the wave-cooperative / lane-serial tail; 8 and 8 by default),
   "t1_coop_max" (largest
   type-1 class, in primitive pairs, that takes the cooperative tail),
   "t1_tail_grid" (most workgroups of a tail launch), "fuse" (1: one launch
   per stage covering every class, with assembly overlapping the large-grid
   fallbacks; 0: per-class launches on forked streams; 2, the default: fused
   for small jobs), "asm2_cache" (type-2 assembly with the angular factors
   in registers: 0 no, 1, the default: yes, 2 with fused launches only),
   "t2_head" (PS93 levels of the type-2 small-grid kernel before the entries
   still open are rerun compacted; 15, the default: no split; 0: 6 with
   fused launches, no split otherwise), "t2_tail_grid",
   "fb2_split" (waves sharing one (triplet, primitive pair) of the type-2
   large-grid fallback; 0, the default: 3 for jobs of up to a million
   (triplet, pair) items, 1 for larger ones).  "t1_spec" (fused launches: 1, the default, starts
   the type-1 large-grid fallback beside the small-grid tail on every entry
   the head left open; 0 after the tail on the failed entries).  "t2_spec"
   (fused launches with a type-2 head: 1 starts the type-2 large-grid
   fallback beside the tail on every entry the head left open; 0, the
   default, after the tail on the failed entries).  Results do not depend on
   them.  0, or -1 for an unknown name. */
int libecp_batch_set_param(libecp_batch *b, const char *name, int value);
int libecp_batch_get_param(libecp_batch *b, const char *name, int *value);
/* Autotuner: times the batch (graph replays, `repeats` per candidate) over
   the launch parameters by coordinate search, keeps the fastest setting,
   and writes what it tried to report (len bytes).  0, or -1. */
int libecp_batch_autotune(libecp_batch *b, int repeats, char *report, int len);
void libecp_batch_free(libecp_batch *b);

/* ---- serving ------------------------------------------------------------
   A request server with continuous batching.  `slots` worker threads each
   own a device context (streams, staging, resident static tables); a free
   worker takes every queued request (up to max_batch_triplets nominal
   triplets per job, 0 = 4,000,000), screens them on the host into one job
   and admits the job when its device block fits in a pool of hbm_budget
   bytes allocated once at creation.  Jobs of different workers run on the
   device at the same time.  A request's result does not depend on what it
   was batched with.  All calls are thread-safe. */

typedef struct libecp_server libecp_server;

/* Uses the calling thread's current HIP device.  NULL on error. */
libecp_server *libecp_server_create(int slots, size_t hbm_budget, long long max_batch_triplets,
                                    int largeGridOrder, double tolerance, double accuracy);
/* Queues one molecule (arrays are copied; invalid input is rejected here).
   Returns its ticket (>= 0), or -1. */
long long libecp_server_submit(libecp_server *s, const libecp_molecule *m);
/* Waits for a request: 0 finished, 1 timed out (timeout_s < 0: no limit),
   -1 unknown ticket. */
int libecp_server_wait(libecp_server *s, long long ticket, double timeout_s);
/* The next finished request not yet returned by this call or collected, in
   completion order; -1 if none finished within timeout_s. */
long long libecp_server_next(libecp_server *s, double timeout_s);
/* Collects a finished request and releases it: adds its ECP matrix into the
   upper triangle of I (row stride rowdim; I may be NULL to discard), and
   when times != NULL writes seconds since server creation: submitted,
   admitted (device memory granted), finished.  Returns 0, 1 / 2 for a
   failed type-1 / type-2 integration (blocks up to the failure added), or
   -1 (message in libecp_last_error(): device error, a request larger than
   the HBM budget, unknown ticket). */
int libecp_server_result(libecp_server *s, long long ticket, int rowdim, double *I, double *times);
/* Counters, up to n of: submitted, finished, jobs run, requests in jobs,
   most jobs on the device at once, largest job (requests), peak pool bytes,
   pool bytes, queued now, running now.  Returns the number available. */
int libecp_server_stats(libecp_server *s, long long *out, int n);
/* Stops the workers after their current jobs; unfinished requests are
   dropped. */
void libecp_server_free(libecp_server *s);

#ifdef __cplusplus
}
#endif
#endif
