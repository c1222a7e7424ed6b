/* Geometry-independent tables of the ECP integral engine (host side, C11).
 *
 * These are built once per handle (or once per process for the batch and
 * serving paths) and uploaded to HBM.  Their values follow the reference's
 * algorithms with the same arithmetic order, so the device sees tables that
 * are bitwise equal to the reference's:
 *   factorials            util.c:13-56
 *   Cartesian ordering    dimensions.c:17-57
 *   cart2sph / poly2sph   transformations.c:28-87, 146-207
 *   Gauss-Chebyshev grids gc_integrators.c:89-145 (PSM92), 220-283 (PS93),
 *                         301-313 (Krack-Koster map)
 *   Bessel table          bessel.c:19-82
 *   angular integrals     angular_integrals.c:15-101
 * The integrator visit schedules are derived from the loop structure of
 * integrateGC_PS93 / integrateGC_PSM92 (gc_integrators.c:156-217, 38-86).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ecp_host.h"

double *ecp_factorials(int n, int step) {
  double *f = calloc((size_t)n + 2, sizeof(double));
  if (!f) return NULL;
  if (n > 0) f[0] = 1.0;
  if (n > 1) f[1] = 1.0;
  for (int i = 2; i <= n; i++) {
    f[i] = f[i - step] * i;
    if (isinf(f[i])) {
      free(f);
      return NULL;
    }
  }
  return f;
}

int *ecp_cart_order(int am) {
  int *t = calloc(3 * (size_t)ECP_DIM_C(am), sizeof(int));
  if (!t) return NULL;
  for (int l = 0; l <= am; l++) {
    int c = 0;
    for (int i = 0; i <= l; i++)
      for (int j = 0; j <= i; j++) {
        int p = 3 * ECP_IDX_C(l, c++);
        t[p] = l - i;
        t[p + 1] = i - j;
        t[p + 2] = j;
      }
  }
  return t;
}

int *ecp_cart_index(int am, const int *ijk) {
  const int dim = am + 1;
  int *idx = calloc((size_t)dim * dim * dim, sizeof(int));
  if (!idx) return NULL;
  for (int l = 0; l <= am; l++)
    for (int c = 0; c < ECP_DIM_IJK(l); c++) {
      int p = 3 * ECP_IDX_C(l, c);
      idx[(ijk[p] * dim + ijk[p + 1]) * dim + ijk[p + 2]] = p / 3;
    }
  return idx;
}

int ecp_tm_offset(int l) {
  int off = 0;
  for (int k = 0; k < l; k++) off += (2 * k + 1) * ECP_DIM_IJK(k);
  return off;
}

static double nk(int n, int k, const double *fac) {
  return (k >= 0 && k <= n) ? fac[n] / (fac[n - k] * fac[k]) : 0.0;
}

static double sgn_pow(int e) { return pow(-1.0, e); }

double *ecp_cart2sph(int lmax, const int *xyz, const double *fac) {
  double *TM = calloc((size_t)ecp_tm_offset(lmax + 1) + 1, sizeof(double));
  if (!TM) return NULL;
  double *T = TM;
  for (int l = 0; l <= lmax; l++)
    for (int m = -l; m <= l; m++) {
      const int mm = abs(m);
      for (int c = 0; c < ECP_DIM_IJK(l); c++, T++) {
        const int *e = xyz + 3 * ECP_IDX_C(l, c);
        const int lx = e[0], ly = e[1], lz = e[2];
        int j = lx + ly - mm;
        if (j < 0 || j % 2 == 1) {
          *T = 0.0;
          continue;
        }
        j /= 2;
        double s1 = 0.0;
        for (int i = 0; i <= (l - mm) / 2; i++) {
          double s2 = 0.0;
          for (int k = 0; k <= j; k++) {
            double s = 0.0;
            if ((m < 0 && abs(mm - lx) % 2 == 1) || (m > 0 && abs(mm - lx) % 2 == 0))
              s = sgn_pow((mm - lx + 2 * k) / 2) * sqrt(2.0);
            else if (m == 0 && lx % 2 == 0)
              s = sgn_pow(-lx / 2 + k);
            s2 += nk(j, k, fac) * nk(mm, (lx - 2 * k), fac) * s;
          }
          s1 += nk(l, i, fac) * nk(i, j, fac) * pow(-1.0, i) * fac[2 * l - 2 * i] / fac[l - mm - 2 * i] * s2;
        }
        *T = sqrt((fac[2 * lx] * fac[2 * ly] * fac[2 * lz] * fac[l] * fac[l - mm]) /
                  (fac[2 * l] * fac[lx] * fac[ly] * fac[lz] * fac[l + mm])) *
             1 / (pow(2.0, l) * fac[l]) * s1;
      }
    }
  return TM;
}

double *ecp_poly2sph(const double *M, int lmax, const int *xyz, const double *fac2) {
  const int ldim = ECP_DIM_L(lmax);
  double *TM = calloc((size_t)ECP_DIM_C(lmax) * ldim, sizeof(double));
  if (!TM) return NULL;
  for (int l1 = 0; l1 <= lmax; l1++)
    for (int c1 = 0; c1 < ECP_DIM_IJK(l1); c1++) {
      const int *e1 = xyz + 3 * ECP_IDX_C(l1, c1);
      double *T = TM + (size_t)ECP_IDX_C(l1, c1) * ldim;
      const double *TCS = M;
      for (int l2 = 0; l2 <= l1; l2++) {
        const double s1 = 4.0 * M_PI * fac2[2 * l2 + 1];
        for (int m = 0; m < 2 * l2 + 1; m++, T++) {
          double sum = 0.0;
          for (int c2 = 0; c2 < ECP_DIM_IJK(l2); c2++, TCS++) {
            const int *e2 = xyz + 3 * ECP_IDX_C(l2, c2);
            const int lx = e1[0] + e2[0], ly = e1[1] + e2[1], lz = e1[2] + e2[2];
            if (lx % 2 || ly % 2 || lz % 2) continue;
            double s = s1, s2 = 1.0 / fac2[l1 + l2 + 1];
            if (lx > 2) s2 *= fac2[lx - 1];
            if (ly > 2) s2 *= fac2[ly - 1];
            if (lz > 2) s2 *= fac2[lz - 1];
            if (e2[0] > 1) s /= fac2[2 * e2[0] - 1];
            if (e2[1] > 1) s /= fac2[2 * e2[1] - 1];
            if (e2[2] > 1) s /= fac2[2 * e2[2] - 1];
            sum += sqrt(s) * s2 * (*TCS);
          }
          *T = sum;
        }
      }
    }
  return TM;
}

/* ---------------------------------------------------------------------- */

int ecp_grid_psm92(int maxPoints, double *x, double *w) {
  const int order = pow(2, floor(log(maxPoints + 1) / log(2))) - 1;
  int offset = (int)pow(2, (int)floor(log(order) / log(2)));
  const int M = (order - 1) / 2;
  int n = 1;
  double N = n + 1.0, S0 = 1.0, C0 = 0.0;
  if (order > ECP_LARGE_MAX || order < 3) return -1;
  memset(x, 0, sizeof(double) * (size_t)order);
  memset(w, 0, sizeof(double) * (size_t)order);
  x[M] = 0.0;
  w[M] = 1.0;
  while (n <= M) {
    const double C1 = C0, S1 = S0;
    C0 = sqrt((1 + C1) / 2);
    S0 = S1 / (2 * C0);
    double s = S0, c = C0;
    offset /= 2;
    for (int i = 1; i <= n; i += 2) {
      double t = 1 + 2 / (3 * M_PI) * (3 + 2 * s * s) * s * c - i / N;
      const int idx = i * offset - 1;
      x[order - idx - 1] = t;
      x[idx] = -t;
      w[order - idx - 1] = w[idx] = s * s * s * s;
      t = s;
      s = s * C1 + c * S1;
      c = c * C1 - t * S1;
    }
    n = 2 * n + 1;
    N = n + 1.0;
  }
  return order;
}

int ecp_grid_ps93_kk(double *x, double *w) {
  const int maxPoints = 128;
  int offset = (int)pow(2, (int)floor(log(maxPoints) / log(2)));
  const int order = 3 * offset - 1;
  int n = 3;
  double C0 = sin(M_PI / 3), S0 = 0.5, C1 = S0, S1 = C0;
  double c = cos(M_PI / 3), s = C0, s2 = s * s, t;
  memset(x, 0, sizeof(double) * (size_t)order);
  memset(w, 0, sizeof(double) * (size_t)order);
  x[order / 2] = 0.0;
  w[order / 2] = 1.0;
  t = (n - 2.0) / n + 2 / M_PI * (1 + 2 * s2 / 3) * c * s;
  x[offset - 1] = -t;
  x[order - offset] = t;
  w[order - offset] = w[offset - 1] = s2 * s2;
  while ((4 * n / 3 - 1) <= order) {
    c = C0;
    s = S0;
    offset /= 2;
    for (int i = 1; i < n; i += 2) {
      s2 = s * s;
      const int idx = i * offset - 1;
      t = 1 + 2 / (3 * M_PI) * s * c * (3 + 2 * s2) - ((double)i) / n;
      x[idx] = -t;
      x[order - idx - 1] = t;
      w[order - idx - 1] = w[idx] = s2 * s2;
      t = s;
      s = s * C1 + c * S1;
      c = c * C1 - t * S1;
    }
    n *= 2;
    C1 = C0;
    S1 = S0;
    C0 = sqrt((1 + C0) / 2);
    S0 = S0 / (2 * C0);
  }
  /* Krack-Koster map onto (0, inf) */
  const double ln2 = log(2.0);
  for (int i = 0; i < order; i++) {
    const double xi = 1.0 - log(1.0 - x[i]) / ln2;
    const double wi = w[i] / (ln2 * (1.0 - x[i]));
    x[i] = xi;
    w[i] = wi;
  }
  return order;
}

/* Visit order of integrateGC_PS93 for the 383-point grid. */
void ecp_schedule_ps93(ecp_schedule *S, int *pairs_l, int *pairs_r) {
  int offset = 128;
  const int order = 3 * offset - 1;
  int n = 3, j = 0, np = 0, lv = 2;
  memset(S, 0, sizeof(*S));
  S->order = order;
  S->centre = (order - 1) / 2;
  S->first_l = offset - 1;
  S->first_r = order - offset;
  offset /= 2;
  while ((2 * n * (1 - j) + j * 4 * n / 3 - 1) <= order) {
    j = 1 - j;
    if (0 == j) offset /= 2;
    S->lvl_begin[lv] = np;
    for (int i = 1; i < n; i += 2)
      if (3 * ((i + 2 * j) / 3) >= i + j) {
        const int idx = i * offset - 1;
        pairs_l[np] = idx;
        pairs_r[np] = order - idx - 1;
        np++;
      }
    n *= (1 + j);
    S->lvl_n[lv] = n;
    S->lvl_j[lv] = j;
    lv++;
  }
  S->lvl_begin[lv] = np;
  S->nlevels = lv;
  S->npairs = np;
}

/* Visit order of integrateGC_PSM92 for a 2^k-1 point grid. */
void ecp_schedule_psm92(ecp_schedule *S, int order, int *pairs_l, int *pairs_r) {
  int offset = (int)pow(2, (int)floor(log(order) / log(2)));
  const int M = (order - 1) / 2;
  int n = 1, np = 0, lv = 1;
  memset(S, 0, sizeof(*S));
  S->order = order;
  S->centre = M;
  while (n <= M) {
    offset /= 2;
    S->lvl_begin[lv] = np;
    for (int i = 1; i <= n; i += 2) {
      const int idx = i * offset - 1;
      pairs_l[np] = idx;
      pairs_r[np] = order - idx - 1;
      np++;
    }
    n = 2 * n + 1;
    S->lvl_n[lv] = n + 1; /* N = n + 1 after the level */
    lv++;
  }
  S->lvl_begin[lv] = np;
  S->nlevels = lv;
  S->npairs = np;
}

/* ---------------------------------------------------------------------- */

/* K_l(z) = exp(-z) i_l(z) on z = i/100, i = 0..N, rows 0..rows-1, written in
   the device layout K[i * ECP_BESSEL_STRIDE + l].  The series of K_0 decides
   the number of terms used for every row, as in the reference. */
int ecp_bessel_table(double accuracy, int N, int cutoff, int rows, double *Kdev) {
  const int lMax = rows - 1;
  double *F = calloc((size_t)cutoff + 2, sizeof(double));
  double *G = calloc((size_t)cutoff + lMax + 3, sizeof(double));
  int rc = 0;
  if (!F || !G) {
    free(F);
    free(G);
    return -1;
  }
  memset(Kdev, 0, sizeof(double) * (size_t)(N + 1) * ECP_BESSEL_STRIDE);
  Kdev[0] = 1.0;
  for (int i = 1; i <= N && !rc; i++) {
    const double z = i / (N / 16.0);
    double f = z * z / 2.0, s, acc0 = 0.0;
    int j = 0, l;
    F[0] = exp(-z);
    G[0] = 1.0;
    s = F[0] / G[0];
    l = (int)(0.25 * sqrt(1.0 + 16.0 * f));
    while (s > accuracy || j <= l) {
      acc0 += s;
      j++;
      if (j > cutoff) {
        rc = 1;
        break;
      }
      F[j] = F[j - 1] * f / j;
      G[j] = G[j - 1] * (2.0 * j + 1.0);
      s = F[j] / G[j];
    }
    if (rc) break;
    Kdev[(size_t)i * ECP_BESSEL_STRIDE] = acc0;
    for (l = 1; l <= lMax; l++) G[j + l] = G[j + l - 1] * (2 * j + 2 * l + 1);
    f = z;
    for (l = 1; l <= lMax; l++) {
      s = 0;
      for (int m = 0; m < j; m++) s += F[m] / G[l + m];
      Kdev[(size_t)i * ECP_BESSEL_STRIDE + l] = f * s;
      f *= z;
    }
  }
  free(F);
  free(G);
  return rc;
}

/* Omega[LM(l,m)][LM(lambda,mu)][C(alpha,c)] for l < maxL, lambda <= maxLambda,
   alpha <= maxAlpha; zero where the reference leaves zeros. */
double *ecp_omega(int maxL, int maxLambda, int maxAlpha, const int *ijk, const int *ijkIndex, int ijkDim,
                  const double *cart2sph, const double *poly2sph, int tmDim, const double *dfac) {
  const int d1 = ECP_DIM_L(maxLambda), d2 = ECP_DIM_C(maxAlpha);
  const int tmd = ECP_DIM_L(tmDim);
  double *O = calloc((size_t)ECP_DIM_L(maxL) * d1 * d2 + 1, sizeof(double));
  if (!O) return NULL;
  for (int lambda = 0; lambda <= maxLambda; lambda++)
    for (int l = 0; l < maxL; l++) {
      const int par = (lambda + l) % 2, lo = lambda - l;
      for (int alpha = par > lo ? par : lo; alpha <= maxAlpha; alpha += 2)
        for (int mu = 0; mu < 2 * lambda + 1; mu++)
          for (int m = 0; m < 2 * l + 1; m++) {
            double *cijk = O + ((size_t)ECP_IDX_LM(l, m) * d1 + ECP_IDX_LM(lambda, mu)) * d2 + ECP_DIM_C(alpha - 1);
            for (int c = 0; c < ECP_DIM_IJK(alpha); c++) {
              const int *e = ijk + 3 * ECP_IDX_C(alpha, c);
              if (0 == alpha) {
                if (l == lambda && m == mu) cijk[c] = 1.0;
                continue;
              }
              if (lambda > l + alpha) continue;
              for (int d = 0; d < ECP_DIM_IJK(l); d++) {
                const int *f = ijk + 3 * ECP_IDX_C(l, d);
                double N = 0.25 * dfac[2 * l + 1] / M_PI;
                if (f[0] > 1) N /= dfac[2 * f[0] - 1];
                if (f[1] > 1) N /= dfac[2 * f[1] - 1];
                if (f[2] > 1) N /= dfac[2 * f[2] - 1];
                N = sqrt(N);
                const int ax = f[0] + e[0], ay = f[1] + e[1], az = f[2] + e[2];
                const int p = ijkIndex[(ax * ijkDim + ay) * ijkDim + az] * tmd + ECP_IDX_LM(lambda, mu);
                const int ti = (l == 0) ? 0 : ecp_tm_offset(l) + m * ECP_DIM_IJK(l) + d;
                cijk[c] += N * cart2sph[ti] * poly2sph[p];
              }
            }
          }
    }
  return O;
}
