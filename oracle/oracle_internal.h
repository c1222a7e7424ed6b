/* Internal declarations of the CPU oracle (test infrastructure only). */
#ifndef ECP_ORACLE_INTERNAL_H
#define ECP_ORACLE_INTERNAL_H 1

#include "oracle.h"

/* Index arithmetic of the reference's dimensions.h:18-29. */
#define ODIM_L(l) (((l) + 1) * ((l) + 1))
#define ODIM_C(l) (((l) + 1) * ((l) + 2) * ((l) + 3) / 6)
#define ODIM_IJK(l) (((l) + 1) * ((l) + 2) / 2)
#define OIDX_LM(l, m) ((l) * (l) + (m))
#define OIDX_C(l, c) (ODIM_C((l) - 1) + (c))

typedef struct {
  int order;      /* number of abscissae */
  int n;          /* size parameter the integrator re-derives its levels from */
  double *x, *w;
  double tol;
  double I;       /* result of the last successful integration */
  int start, end; /* window; both ends inclusive for the integrator */
} orc_grid;

typedef double (*orc_integrand)(int index, const void *params);

typedef struct {
  int lMax, N;
  int dimK;
  double *K;      /* (lMax+1) x (N+1), row-major in l */
  double *C;      /* recurrence factors l/(2l+1) */
} orc_bessel;

typedef struct {
  int l;
  double a, d, n;
} orc_gauss;

typedef struct {
  int L;          /* local channel label = highest l on the atom */
  int N;          /* number of Gaussians */
  orc_gauss *g;
} orc_ecp;

/* Op-counting build (liboracle_ops.so, compiled as C++ with a counting
   double): the phase the arithmetic is attributed to.  A no-op otherwise. */
enum {
  ORC_PH_OTHER = 0, ORC_PH_UTAB, ORC_PH_FTAB, ORC_PH_ANG, ORC_PH_T1, ORC_PH_FB1, ORC_PH_ASM1,
  ORC_PH_T2, ORC_PH_FB2, ORC_PH_ASM2, ORC_PH_SHIFT, ORC_PH_N
};
#ifdef ORC_OPCOUNT
extern int orc_phase;
#define ORC_PHASE(p) (orc_phase = (p))
#else
#define ORC_PHASE(p) ((void)0)
#endif

/* tables */
double *orc_fac_table(int n, int step);
int *orc_cart_order(int am);
int *orc_cart_order_index(int am, const int *ijk);
double *orc_cart2sph(int lmax, const int *ijk, const double *fac);
double *orc_poly2sph(const double *M, int lmax, const int *ijk, const double *dfac);
int orc_tm_offset(int l);  /* start of the l block of cart2sph */

orc_grid *orc_grid_new_ps93(int maxPoints, double tol);
orc_grid *orc_grid_new_psm92(int maxPoints, double tol);
orc_grid *orc_grid_copy(const orc_grid *g);
void orc_grid_free(orc_grid *g);
void orc_map_kk(orc_grid *g);
void orc_map_fm06(orc_grid *g, double zeta_p, double P);
int orc_integrate_ps93(orc_integrand f, const void *params, orc_grid *t);
int orc_integrate_psm92(orc_integrand f, const void *params, orc_grid *t);

int orc_bessel_init(orc_bessel *b, int lMax, int N, int cutoff, double accuracy);
void orc_bessel_release(orc_bessel *b);
void orc_bessel_weighted(const orc_bessel *b, int dim, int lmax, double z, double *K);

void orc_rsh(int lmax, double theta, double phi, const double *fac, const double *dfac, double *out);
void orc_sph_coords(const double *xyz, double *rtp);

double orc_eval_ecp(const orc_ecp *U, int l, double r);
double orc_gto_radius(double c, double zeta, int l, double cutoff, int *ok);
double orc_shell_radius_(int K, int l, const double *d, const double *a, double zero, int *ok);

#endif
