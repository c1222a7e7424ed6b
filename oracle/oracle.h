/* CPU oracle for the ECP-integral hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, in plain C, of the reference library's algorithm for the
 * semi-local ECP integrals (chrr/libECP, the C files under src/).  It follows the reference's
 * arithmetic order so that, on the same inputs, it reproduces the reference's
 * numbers bit for bit (checked by tests against oracle/_ref, the reference
 * compiled from its own sources).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, as a checker; the product library
 * never links or calls it.
 *
 * Differences from the reference, each behind an option:
 *  - error paths: the reference's failure paths free and then reuse arrays
 *    (type1.c:189-190, type2.c:573-575, libecp.c:358); the oracle instead
 *    returns the codes documented in libecp.h:23-26 (1 = type-1 failure,
 *    2 = type-2 failure) after delivering every block that precedes the
 *    failing one in the reference's callback order.
 *  - ORC_OPT_CLEAN_FALLBACK: in the large-grid fallback the reference keeps
 *    integrand buffers from earlier work (type-1: the small-grid tabulation,
 *    type1.c:171-175; type-2: the previous primitive pair, type2.c:443-448)
 *    and the PSM92 sequence still visits left-half points past a screening
 *    break (gc_integrators.c:63).  With this option those points contribute
 *    zero, which is the semantics of the device library.
 */
#ifndef ECP_ORACLE_H
#define ECP_ORACLE_H 1

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*orc_callback)(int A, int sa, int la, int shifta,
                             int B, int sb, int lb, int shiftb,
                             int C, double *block, void *args);

enum {
  ORC_OPT_CLEAN_FALLBACK = 1,
  /* sensitivity probe: scale every type-1 small-grid integrand value by
     (1 + 2^-50), i.e. a few ulp, to expose convergence decisions that a
     last-bit change flips */
  ORC_OPT_PERTURB_T1 = 2,
};

typedef struct {
  long ps93_calls[2], ps93_fail[2];       /* [type-1, type-2] */
  long psm92_calls[2], psm92_fail[2];
  long left_breaks[2];
  long blocks[2];
  long triplets_evaluated;                /* (C, a, b) with a non-empty window */
  long fb_tabulations[2];                 /* primitive pairs tabulated on the large grid */
  long fb_points[2];                      /* large-grid points with an evaluated integrand */
  long ps93_level[2][16];                 /* convergence level histogram (15 = failed) */
  long psm92_level[2][11];                /* convergence level histogram (10 = failed) */
} orc_stats;

typedef struct orc_handle orc_handle;

/* Same positional meaning as the reference's libECP_init definition
   (libecp.c:54-61): argument 4 is lECP, argument 5 is KECP. */
orc_handle *orc_init(int nrAtoms, const double *geometry,
                     const int *shellsECP, const int *lECP, const int *KECP,
                     const double *nECP, const double *dECP, const double *aECP,
                     const int *shellsBS, const int *lBS, const int *KBS,
                     const double *dBS, const double *aBS,
                     int n, int lmax, const int *shellOrdering,
                     int largeGridOrder, double tolerance, double accuracy,
                     int options);
int orc_calculate(orc_handle *h, orc_callback cb, void *args);
void orc_free(orc_handle *h);
void orc_get_stats(const orc_handle *h, orc_stats *out);

/* One-call form, argument order of getIntegrals.h:7-13 (KECP before lECP). */
int orc_getIntegrals(int nrAtoms, const double *geometry,
                     const int *shellsECP, const int *KECP, const int *lECP,
                     const double *nECP, const double *dECP, const double *aECP,
                     const int *shellsBS, const int *lBS, const int *KBS,
                     const double *dBS, const double *aBS,
                     int largeGridOrder, double tolerance, double accuracy,
                     int rowdim, double *I, int options);

/* Op-counting build only (liboracle_ops.so): FP64 operations executed,
   per phase of the algorithm (ORC_PH_* in oracle_internal.h), since the
   last reset. */
void orc_ops_reset(void);
int orc_ops_get(long long *out, int n);

/* Table-level entry points used by the unit tests. */
void orc_grid_ps93(int maxPoints, double *x, double *w, int *order);   /* KK-mapped */
void orc_grid_psm92(int maxPoints, double *x, double *w, int *order);  /* on (-1,1) */
int orc_bessel_table(int lMax, int N, int cutoff, double accuracy, double *K /* (lMax+1)*(N+1) */);
void orc_bessel_eval(int lMax, double accuracy, int lmax, double z, double *K /* lmax+1 */);
void orc_real_sph(int lmax, double theta, double phi, double *out /* (lmax+1)^2 */);
void orc_spherical_coordinates(const double *xyz, double *rtp);
double orc_quad_ps93(double (*f)(double), double tol, int *rc);
double orc_quad_psm92(int maxPoints, double zeta_p, double P, double (*f)(double), double tol, int *rc);
double orc_shell_radius(int K, int l, const double *d, const double *a, double zero);

#ifdef __cplusplus
}
#endif
#endif
