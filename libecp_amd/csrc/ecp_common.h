/* Shared definitions of the host library (C) and the HIP kernels.
 *
 * Everything here is plain C: integer index arithmetic of the Cartesian /
 * spherical layouts (the reference's dimensions.h:18-29 conventions), the
 * integrator visit schedules, and the flat job description the host builds
 * and the device consumes.
 */
#ifndef ECP_COMMON_H
#define ECP_COMMON_H 1

#include <stdint.h>

#define ECP_DIM_L(l) (((l) + 1) * ((l) + 1))
#define ECP_DIM_C(l) (((l) + 1) * ((l) + 2) * ((l) + 3) / 6)
#define ECP_DIM_IJK(l) (((l) + 1) * ((l) + 2) / 2)
#define ECP_IDX_LM(l, m) ((l) * (l) + (m))
#define ECP_IDX_C(l, c) (ECP_DIM_C((l) - 1) + (c))

/* Small grid: PS93 with 128 "max points" gives 3*128-1 = 383 abscissae
   (type1.c:24, type2.c:60). */
#define ECP_SMALL_ORDER 383
#define ECP_SMALL_LEVELS 15
/* Largest supported large-grid order (PSM92: 2^k - 1). */
#define ECP_LARGE_MAX 2047
#define ECP_LARGE_LEVELS_MAX 11

/* Limits of this build.  Basis l plus derivative order <= 6 (type-1
   classes la + lb <= 12, Bessel orders <= 12), ECP local label L <= 7.  The Bessel table carries rows 0..ECP_BESSEL_ROWS-1. */
#define ECP_MAX_LBS 6
#define ECP_MAX_L 7
#define ECP_MAX_LAB (2 * ECP_MAX_LBS)
#define ECP_BESSEL_ROWS 24
#define ECP_BESSEL_N 1600
#define ECP_BESSEL_STRIDE 24 /* doubles per table abscissa in the device layout */

/* Integrator visit schedule.  Level 0 is the centre point, level 1 the first
   pair (both evaluated without a window test); every later level is a list
   of (left, right) point pairs visited in order, each point tested against
   the window (left: index >= start, right: index <= end). */
typedef struct {
  int order;                 /* number of abscissae */
  int nlevels;               /* levels including 0 and 1 */
  int centre;                /* level-0 point */
  int first_l, first_r;      /* PS93 level-1 pair (unused for PSM92) */
  int lvl_begin[ECP_LARGE_LEVELS_MAX + ECP_SMALL_LEVELS + 2]; /* into pairs[] */
  int lvl_n[ECP_LARGE_LEVELS_MAX + ECP_SMALL_LEVELS + 2];     /* PS93: n after the level; PSM92: N */
  int lvl_j[ECP_SMALL_LEVELS + 2];                            /* PS93 one/two-point switch */
  int npairs;
} ecp_schedule;

/* Functions shared by the host model and the device kernels (plain C in
   the host library, host + device functions in the HIP source). */
#ifdef __HIPCC__
#define ECP_HD __host__ __device__
#else
#define ECP_HD
#endif

/* FLOP model of the large-grid fallbacks (the reference's eager algorithm,
   type1.c:148-196, type2.c:417-528; FMA = 2, libm calls 1):
   type 1, a primitive pair of class lab with f failed small-grid entries:
   N points tabulated (exponent, Bessel orders 0..lab, U_L, powers of r)
   and f integrals over them;
   type 2, one primitive pair of a (triplet, channel l) group with f failed
   entries: N points with two Bessel vectors (orders to la + l and lb + l),
   the exponent and its exp, U_l (ng Gaussians), the powers r^k, and per
   entry and point the five-factor integrand, its weight and sum. */
static inline ECP_HD double ecp_fb1_flops(int lab, int f, int N) {
  return f > 0 ? (double)N * (31.0 * lab + 74) + f * (4.0 * N + 150) : 0.0;
}
static inline ECP_HD double ecp_fb2_flops(int la, int lb, int l, int ng, int f, int N) {
  return f > 0 ? (double)N * (30.0 * (2 * l + la + lb) + 144 + 5.0 * ng + la + lb) + f * (8.0 * N + 150) : 0.0;
}

/* ---- flat job description (host -> device) ---------------------------- */

typedef struct {
  double x, y, z;
} ecp_xyz;

typedef struct {
  int atom;      /* global atom index */
  int l;
  int K;         /* primitives */
  int prim;      /* offset into prims */
  int mol;
  int local;     /* shell index inside its molecule */
} ecp_shell;

typedef struct {
  int atom;      /* global atom index of the ECP centre */
  int L;         /* local channel label; semi-local channels 0..L-1 */
  int g0, ng;    /* Gaussians [g0, g0+ng) */
  int mol;
  int end[ECP_MAX_L + 1];  /* progressive potential cut-off per semi-local channel */
  int UL_off;    /* offset (doubles) of the U_L table in ULtab */
  int Ut_off;    /* offset of Ut[l][k][383] in Uttab */
  int F_off;     /* offset of this centre's F slab block in Ftab */
  int win_off;   /* offset into windows (per shell of the molecule) */
  int atom0, nat;   /* the molecule's atoms */
  int shell0, nsh;  /* the molecule's shells */
  int ang_off;   /* offset of the per-(centre, atom) angular tables */
} ecp_centre;

typedef struct {
  int l;
  double n, d, a;
} ecp_gauss;

typedef struct {
  int start, end;  /* screened window of the small grid, end inclusive in the
                      reference's sense (-1: empty) */
  int skip;
} ecp_window;

typedef struct {
  int centre;      /* global centre index */
  int sa, sb;      /* global shell indices */
  int gs, ge;      /* pair window: [max start, max end] */
  int blk1, blk2;  /* output offsets (doubles) of the type-1 / type-2 blocks */
  int la, lb;      /* angular momenta of the integrals: shell l + shift (derivatives) */
  int oa, ob;      /* angular momenta of the output block (la, lb, except the unshifted
                      side of the second-derivative shifts (+1, 0) and (0, +1)) */
  int powa, powb;  /* contraction coefficients are scaled by zeta^pow (positive shifts) */
  int shifta, shiftb;
} ecp_triplet;

#endif
