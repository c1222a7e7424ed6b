/* CPU oracle, part 1: tables and numerical primitives (TEST INFRASTRUCTURE).
 *
 * Each routine restates the reference's algorithm with the same arithmetic
 * order, so that results agree bit for bit with the reference built from its
 * sources.  Citations are to the reference's src/ directory.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

/* n! (step 1) or n!! (step 2) for 0..n; util.c:13-56.  The reference aborts
   when a value overflows; the oracle returns NULL instead. */
double *orc_fac_table(int n, int step) {
  double *f = calloc((size_t)n + 1, sizeof(double));
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

/* Cartesian exponents in libint order for l = 0..am; dimensions.c:17-38. */
int *orc_cart_order(int am) {
  int *t = calloc(3 * (size_t)ODIM_C(am), sizeof(int));
  for (int l = 0; l <= am; l++) {
    int c = 0;
    for (int i = 0; i <= l; i++) {
      for (int j = 0; j <= i; j++) {
        int p = 3 * OIDX_C(l, c);
        t[p] = l - i;
        t[p + 1] = i - j;
        t[p + 2] = j;
        c++;
      }
    }
  }
  return t;
}

/* (i,j,k) -> position in the ordering table; dimensions.c:41-57. */
int *orc_cart_order_index(int am, const int *ijk) {
  const int dim = am + 1;
  int *idx = calloc((size_t)dim * dim * dim, sizeof(int));
  for (int l = 0; l <= am; l++) {
    for (int c = 0; c < ODIM_IJK(l); c++) {
      int p = 3 * OIDX_C(l, c);
      idx[ijk[p] * dim * dim + ijk[p + 1] * dim + ijk[p + 2]] = p / 3;
    }
  }
  return idx;
}

int orc_tm_offset(int l) {
  int off = 0;
  for (int k = 0; k < l; k++) off += (2 * k + 1) * ODIM_IJK(k);
  return off;
}

static double orc_binom(int n, int k, const double *fac) {
  return (k >= 0 && k <= n) ? fac[n] / (fac[n - k] * fac[k]) : (double)0.0;
}

/* Real-solid-harmonic coefficients of the Cartesian monomials, Schlegel &
   Frisch (1995) eq. 15; transformations.c:28-87.  Layout: for each l, for
   m = -l..l, the IJK_DIM(l) Cartesian coefficients. */
double *orc_cart2sph(int lmax, const int *xyz, const double *fac) {
  double *TM = calloc((size_t)orc_tm_offset(lmax + 1), sizeof(double));
  double *T = TM;
  for (int l = 0; l <= lmax; l++) {
    for (int m = -l; m <= l; m++) {
      int mm = abs(m);
      for (int c = 0; c < ODIM_IJK(l); c++) {
        int p = 3 * OIDX_C(l, c);
        int lx = xyz[p], ly = xyz[p + 1], lz = xyz[p + 2];
        int j = lx + ly - mm;
        if (j < 0 || j % 2 == 1) {
          *T = 0.0;
        } else {
          double s1 = 0.0;
          j = j / 2;
          for (int i = 0; i <= (l - mm) / 2; i++) {
            double s2 = 0.0;
            for (int k = 0; k <= j; k++) {
              double s = 0.0;
              if ((m < 0 && abs(mm - lx) % 2 == 1) || (m > 0 && abs(mm - lx) % 2 == 0)) {
                int e = (mm - lx + 2 * k) / 2;
                s = pow(-1.0, e) * sqrt(2.0);
              } else if (m == 0 && lx % 2 == 0) {
                int e = -lx / 2 + k;
                s = pow(-1.0, e);
              }
              s2 += orc_binom(j, k, fac) * orc_binom(mm, (lx - 2 * k), fac) * s;
            }
            s1 += orc_binom(l, i, fac) * orc_binom(i, j, fac) * pow(-1.0, i) *
                  fac[2 * l - 2 * i] / fac[l - mm - 2 * i] * s2;
          }
          *T = sqrt((fac[2 * lx] * fac[2 * ly] * fac[2 * lz] * fac[l] * fac[l - mm]) /
                    (fac[2 * l] * fac[lx] * fac[ly] * fac[lz] * fac[l + mm])) *
               1 / (pow(2.0, l) * fac[l]) * s1;
        }
        T++;
      }
    }
  }
  return TM;
}

/* Expansion of unit-sphere monomials in real spherical harmonics, Flores-
   Moreno et al. (2006) eq. 36; transformations.c:146-207.  Row = Cartesian
   monomial (C_DIM(lmax) rows), column = LM index (L_DIM(lmax) columns). */
double *orc_poly2sph(const double *M, int lmax, const int *xyz, const double *fac2) {
  const int ldim = ODIM_L(lmax);
  double *TM = calloc((size_t)ODIM_C(lmax) * ldim, sizeof(double));
  for (int l1 = 0; l1 <= lmax; l1++) {
    for (int c1 = 0; c1 < ODIM_IJK(l1); c1++) {
      int p = 3 * OIDX_C(l1, c1);
      int lx1 = xyz[p], ly1 = xyz[p + 1], lz1 = xyz[p + 2];
      double *T = TM + (size_t)OIDX_C(l1, c1) * ldim;
      const double *TCS = M;
      for (int l2 = 0; l2 <= l1; l2++) {
        double s1 = 4.0 * M_PI * fac2[2 * l2 + 1];
        for (int m = 0; m < 2 * l2 + 1; m++) {
          double sum = 0.0;
          for (int c2 = 0; c2 < ODIM_IJK(l2); c2++) {
            int q = 3 * OIDX_C(l2, c2);
            int lx2 = xyz[q], ly2 = xyz[q + 1], lz2 = xyz[q + 2];
            int lx = lx1 + lx2, ly = ly1 + ly2, lz = lz1 + lz2;
            if (lx % 2 == 0 && ly % 2 == 0 && lz % 2 == 0) {
              double s = s1;
              double s2 = 1.0 / fac2[l1 + l2 + 1];
              if (lx > 2) s2 *= fac2[lx - 1];
              if (ly > 2) s2 *= fac2[ly - 1];
              if (lz > 2) s2 *= fac2[lz - 1];
              if (lx2 > 1) s /= fac2[2 * lx2 - 1];
              if (ly2 > 1) s /= fac2[2 * ly2 - 1];
              if (lz2 > 1) s /= fac2[2 * lz2 - 1];
              sum += sqrt(s) * s2 * (*TCS);
            }
            TCS++;
          }
          *T = sum;
          T++;
        }
      }
    }
  }
  return TM;
}

/* ---------------------------------------------------------------------- */
/* Gauss-Chebyshev (2nd kind) grids and adaptive integrators.             */

static orc_grid *grid_alloc(int order) {
  orc_grid *g = calloc(1, sizeof(orc_grid));
  g->order = order;
  g->x = calloc((size_t)order, sizeof(double));
  g->w = calloc((size_t)order, sizeof(double));
  g->start = 0;
  g->end = order - 1;
  return g;
}

orc_grid *orc_grid_copy(const orc_grid *o) {
  orc_grid *g = grid_alloc(o->order);
  g->n = o->n;
  g->tol = o->tol;
  g->I = o->I;
  g->start = o->start;
  g->end = o->end;
  memcpy(g->x, o->x, sizeof(double) * (size_t)o->order);
  memcpy(g->w, o->w, sizeof(double) * (size_t)o->order);
  return g;
}

void orc_grid_free(orc_grid *g) {
  if (!g) return;
  free(g->x);
  free(g->w);
  free(g);
}

/* Nested 2^k-1 point grid of Perez-Jorda, San-Fabian & Moscardo (1992);
   gc_integrators.c:89-145. */
orc_grid *orc_grid_new_psm92(int maxPoints, double tol) {
  int order = pow(2, floor(log(maxPoints + 1) / log(2))) - 1;
  int runs = (int)floor(log(order) / log(2));
  int offset = (int)pow(2, runs);
  int M = (order - 1) / 2;
  int n = 1;
  double N = n + 1.0;
  double S0 = 1.0, C0 = 0.0, S1, C1, s, c, t;
  orc_grid *T = grid_alloc(order);
  double *x = T->x, *w = T->w;
  T->n = order;
  T->tol = tol;
  x[M] = 0.0;
  w[M] = 1.0;
  while (n <= M) {
    C1 = C0;
    S1 = S0;
    C0 = sqrt((1 + C1) / 2);
    S0 = S1 / (2 * C0);
    s = S0;
    c = C0;
    offset /= 2;
    for (int i = 1; i <= n; i += 2) {
      t = 1 + 2 / (3 * M_PI) * (3 + 2 * s * s) * s * c - i / N;
      int idx = i * offset - 1;
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
  return T;
}

/* 3*2^k-1 point grid of Perez-Jorda & San-Fabian (1993);
   gc_integrators.c:220-283. */
orc_grid *orc_grid_new_ps93(int maxPoints, double tol) {
  int runs = (int)floor(log(maxPoints) / log(2));
  int offset = (int)pow(2, runs);
  int order = 3 * offset - 1;
  orc_grid *T = grid_alloc(order);
  double *x = T->x, *w = T->w;
  double C0, S0, S1, C1, s, s2, c, t;
  int n = 3;
  T->n = maxPoints;
  T->tol = tol;
  C0 = sin(M_PI / 3);
  S0 = 0.5;
  C1 = S0;
  S1 = C0;
  c = cos(M_PI / 3);
  s = C0;
  s2 = s * s;
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
      int idx = i * offset - 1;
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
  return T;
}

/* Krack-Koster map of (-1,1) onto (0,inf): r = 1 - log2(1-x);
   gc_integrators.c:301-313. */
void orc_map_kk(orc_grid *g) {
  double ln2 = log(2.0);
  for (int i = 0; i < g->order; i++) {
    double xi = 1.0 - log(1.0 - g->x[i]) / ln2;
    double wi = g->w[i] / (ln2 * (1.0 - g->x[i]));
    g->x[i] = xi;
    g->w[i] = wi;
  }
}

/* Linear map onto [max(0, P-7 sigma), P+9 sigma], sigma = zeta_p^-1/2;
   gc_integrators.c:316-331. */
void orc_map_fm06(orc_grid *g, double zeta_P, double P) {
  double sigma = 1.0 / sqrt(zeta_P);
  double t = P - 7.0 * sigma;
  double rmin = (t > 0.0) ? t : (double)0.0;
  double rmax = P + 9.0 * sigma;
  double i1 = 0.5 * (rmax - rmin);
  double i2 = 0.5 * (rmax + rmin);
  for (int i = 0; i < g->order; i++) {
    g->x[i] = i1 * g->x[i] + i2;
    g->w[i] *= i1;
  }
}

/* Point-doubling integrator with the criterion 16 e^2 <= 3 N |I - q| tol;
   gc_integrators.c:38-86. */
int orc_last_level;
int orc_integrate_psm92(orc_integrand f, const void *ps, orc_grid *t) {
  int level = 0;
  int order = pow(2, floor(log(t->n + 1) / log(2))) - 1;
  const double *w = t->w;
  int runs = (int)floor(log(order) / log(2));
  int offset = (int)pow(2, runs);
  const int M = (order - 1) / 2;
  double I = w[M] * f(M, ps);
  int n = 1;
  double N = n + 1.0;
  double e, T, q, p = I;
  while (n <= M) {
    int cnt = 0;
    q = 2 * p;
    p = 2 * I;
    offset /= 2;
    for (int i = 1; i <= n; i += 2) {
      int idx = i * offset - 1;
      T = 0.0;
      if (idx >= t->start) {
        T += w[idx] * f(idx, ps);
        cnt++;
      }
      if (order - idx - 1 <= t->end) {
        T += w[order - idx - 1] * f(order - idx - 1, ps);
        cnt++;
      }
      I += T;
    }
    n = 2 * n + 1;
    N = n + 1.0;
    e = I - p;
    level++;
    if (0 == cnt) continue;
    if (16 * e * e <= 3 * N * fabs(I - q) * t->tol) {
      t->I = 16 * I / (3 * N);
      orc_last_level = level;
      return 0;
    }
  }
  orc_last_level = 10;
  return 1;
}

/* Interleaved x4/3, x3/2 sequence with the error estimate of PS93;
   gc_integrators.c:156-217. */
int orc_integrate_ps93(orc_integrand f, const void *ps, orc_grid *t) {
  int runs = (int)floor(log(t->n) / log(2));
  int offset = (int)pow(2, runs);
  int order = 3 * offset - 1;
  const double *w = t->w;
  int n = 3, j = 0;
  int idx = offset - 1;
  double p, q, I, err = 0.0, T;
  p = w[(order - 1) / 2] * f((order - 1) / 2, ps);
  q = w[idx] * f(idx, ps) + w[order - offset] * f(order - offset, ps);
  I = p + q;
  offset /= 2;
  int level = 1;
  while ((2 * n * (1 - j) + j * 4 * n / 3 - 1) <= order) {
    int cnt = 0;
    j = 1 - j;
    if (0 == j) offset /= 2;
    for (int i = 1; i < n; i += 2) {
      if (3 * ((i + 2 * j) / 3) >= i + j) {
        idx = i * offset - 1;
        T = 0.0;
        if (idx >= t->start) {
          T += w[idx] * f(idx, ps);
          cnt++;
        }
        if (order - idx - 1 <= t->end) {
          T += w[order - idx - 1] * f(order - idx - 1, ps);
          cnt++;
        }
        I += T;
      }
    }
    n *= (1 + j);
    p += (1 - j) * (I - q);
    if (0 < cnt) err = 16 * fabs((1 - j) * (q - 3 * p / 2) + j * (I - 2 * q)) / (3 * n);
    q = (1 - j) * q + j * I;
    level++;
    if (0 == cnt) continue;
    if (err < t->tol) {
      t->I = 16 * q / (3 * n);
      orc_last_level = level;
      return 0;
    }
  }
  orc_last_level = 15;
  return 1;
}

/* ---------------------------------------------------------------------- */
/* Exponentially weighted modified spherical Bessel functions             */
/* K_l(z) = exp(-z) i_l(z).                                               */

/* Table on z = i/100, i = 0..N, by the power series of K_0 with the term
   count of l = 0 reused for every l; bessel.c:19-82. */
int orc_bessel_init(orc_bessel *b, int lMax, int N, int cutoff, double accuracy) {
  int dim = N + 1;
  double *F = calloc((size_t)cutoff + 1, sizeof(double));
  double *G = calloc((size_t)cutoff + lMax + 2, sizeof(double));
  double *K = calloc((size_t)(lMax + 1) * dim, sizeof(double));
  double *C = calloc((size_t)lMax + 1, sizeof(double));
  b->lMax = lMax;
  b->N = N;
  b->dimK = dim;
  b->K = K;
  b->C = C;
  K[0] = 1.0;
  for (int i = 1; i <= N; i++) {
    double z = i / (N / 16.0);
    int j = 0, l;
    double f = z * z / 2.0, s;
    F[j] = exp(-z);
    G[j] = 1.0;
    s = F[j] / G[j];
    l = (int)(0.25 * sqrt(1.0 + 16.0 * f));
    while (s > accuracy || j <= l) {
      K[i] += s;
      j++;
      if (j > cutoff) {
        free(F);
        free(G);
        return 1;
      }
      F[j] = F[j - 1] * f / j;
      G[j] = G[j - 1] * (2.0 * j + 1.0);
      s = F[j] / G[j];
    }
    for (l = 1; l <= lMax; l++) G[j + l] = G[j + l - 1] * (2 * j + 2 * l + 1);
    f = z;
    for (l = 1; l <= lMax; l++) {
      s = 0;
      for (int m = 0; m < j; m++) s += F[m] / G[l + m];
      K[l * dim + i] = f * s;
      f *= z;
    }
  }
  for (int i = 1; i <= lMax; i++) C[i] = i / (2.0 * i + 1.0);
  free(F);
  free(G);
  return 0;
}

void orc_bessel_release(orc_bessel *b) {
  free(b->K);
  free(b->C);
  b->K = b->C = NULL;
}

/* K_0..K_lmax at z, written with stride dim; bessel.c:102-199:
   z < 1e-7: leading term; z < 16: 5-term Taylor series about the nearest
   table point with derivatives from the recurrence
   K_j' = j/(2j+1) (K_{j-1} - K_{j+1}) - K_j + K_{j+1};
   z >= 16: asymptotic series sum_i (l+i)!/(i!(l-i)!) (-1)^i / (2z)^(i+1). */
void orc_bessel_weighted(const orc_bessel *b, int dim, int lmax, double z, double *K) {
  const double small = 1.0E-7;
  if (z < small) {
    if (z <= 0) {
      K[0] = 1.0;
      for (int l = 1; l <= lmax; l++) K[l * dim] = 0.0;
    } else {
      K[0] = 1 - z;
      for (int l = 1; l <= lmax; l++) K[l * dim] = K[(l - 1) * dim] * z / (2 * l + 1);
    }
  } else if (z < 16.0) {
    int maxLambda = lmax + 5;
    double dKi[64], dKj[64];
    double scale = b->N / 16;
    int index = (int)floor(z * scale + 0.5);
    double dz = z - index / scale;
    const int ddim = b->dimK;
    scale = 1.0;
    for (int l = 0; l <= lmax; l++) K[l * dim] = dKi[l] = b->K[l * ddim + index];
    for (int l = lmax + 1; l <= maxLambda; l++) dKi[l] = b->K[l * ddim + index];
    for (int i = 1; i <= 5; i++) {
      int top = maxLambda - i;
      for (int j = 0; j <= top + 1; j++) dKj[j] = dKi[j];
      dKi[0] = dKj[1] - dKj[0];
      for (int j = 1; j <= top; j++)
        dKi[j] = b->C[j] * (dKj[j - 1] - dKj[j + 1]) - dKj[j] + dKj[j + 1];
      scale = scale * dz / i;
      for (int j = 0; j <= lmax; j++) K[j * dim] += scale * dKi[j];
    }
  } else {
    double A[64];
    A[0] = 0.5 / z;
    for (int l = 0; l <= lmax; l++) K[l * dim] = A[0];
    for (int l = 1; l <= lmax; l++) {
      double f = l * (l + 1);
      for (int i = 1; i < l; i++) {
        K[l * dim] += f * A[i];
        f *= (l + i + 1) * (l - i);
      }
      A[l] = -A[0] * A[l - 1] / l;
      K[l * dim] += f * A[l];
    }
  }
}

/* ---------------------------------------------------------------------- */

/* Real spherical harmonics without the Condon-Shortley phase, normalised
   with sqrt((2l+1)/(2 pi) (l-m)!/(l+m)!) and 1/sqrt(2) for m = 0;
   spherical_harmonics.c:15-114.  out[l*l + l + m] for m = -l..l. */
void orc_rsh(int lmax, double theta, double phi, const double *fac, const double *dfac, double *rsh) {
#define RSH(l, m) ((l) * (l) + (l) + (m))
  double *s = calloc((size_t)lmax + 2, sizeof(double));
  double *c = calloc((size_t)lmax + 2, sizeof(double));
  double *P = calloc((size_t)ODIM_L(lmax), sizeof(double));
  double x = cos(theta);
  if (1.0 == x) {
    for (int l = 0; l <= lmax; l++) P[RSH(l, 0)] = 1.0;
  } else if (-1.0 == x) {
    P[RSH(0, 0)] = 1.0;
    for (int l = 1; l <= lmax; l++) P[RSH(l, 0)] = -P[RSH(l - 1, 0)];
  } else {
    s[1] = sqrt(1.0 - x * x);
    for (int l = 2; l <= lmax; l++) s[l] = s[l - 1] * s[1];
    for (int l = 0; l <= lmax; l++) {
      if (0 == l) {
        P[RSH(l, 0)] = 1.0;
      } else {
        int m = l;
        P[RSH(l, m)] = s[m] * dfac[2 * m - 1];
        m = l - 1;
        P[RSH(l, m)] = x * (2 * m + 1) * P[RSH(l - 1, m)];
        if (l > 1) {
          for (m = 0; m <= l - 2; m++)
            P[RSH(l, m)] = (x * (2 * l - 1) * P[RSH(l - 1, m)] - (l + m - 1) * P[RSH(l - 2, m)]) / (l - m);
        }
      }
    }
  }
  for (int l = 0; l <= lmax; l++) {
    double norm0 = sqrt((2.0 * l + 1.0) / (2.0 * M_PI));
    P[RSH(l, 0)] = norm0 * P[RSH(l, 0)];
    for (int m = 1; m <= l; m++) {
      double norm = sqrt(fac[l - m] / fac[l + m]) * norm0;
      P[RSH(l, m)] = norm * P[RSH(l, m)];
    }
  }
  if (lmax > 0) {
    if (0.0 == phi) {
      for (int m = 0; m <= lmax; m++) {
        s[m] = 0.0;
        c[m] = 1.0;
      }
    } else {
      s[1] = sin(phi);
      c[1] = cos(phi);
      for (int m = 2; m <= lmax; m++) {
        s[m] = s[1] * c[m - 1] + c[1] * s[m - 1];
        c[m] = c[1] * c[m - 1] - s[1] * s[m - 1];
      }
    }
  }
  for (int l = 0; l <= lmax; l++) {
    rsh[RSH(l, 0)] = P[RSH(l, 0)] / sqrt(2.0);
    for (int m = 1; m <= l; m++) {
      rsh[RSH(l, -m)] = P[RSH(l, m)] * s[m];
      rsh[RSH(l, +m)] = P[RSH(l, m)] * c[m];
    }
  }
  free(P);
  free(s);
  free(c);
#undef RSH
}

/* (r, theta, phi) with 1e-14 guards; util.c:74-106. */
void orc_sph_coords(const double *xyz, double *out) {
  const double acc = 1.0E-14;
  const double x = xyz[0], y = xyz[1], z = xyz[2];
  double r = sqrt(x * x + y * y + z * z), theta, phi;
  theta = (r < acc) ? (double)0.0 : acos(z / r);
  if (fabs(x) < acc) {
    if (fabs(y) < acc) phi = 0.0;
    else if (y < 0.0) phi = 1.5 * M_PI;
    else phi = 0.5 * M_PI;
  } else {
    phi = (x > 0.0) ? atan(y / x) : atan(y / x) + M_PI;
  }
  out[0] = r;
  out[1] = theta;
  out[2] = phi;
}

/* U_l(r) = sum_{g: l_g = l} d_g r^n_g exp(-a_g r^2); ecp.c:41-60. */
double orc_eval_ecp(const orc_ecp *U, int l, double r) {
  double potential = 0.0;
  const double r2 = r * r;
  for (int i = 0; i < U->N; i++) {
    const orc_gauss *g = &U->g[i];
    if (l == g->l) potential += pow(r, g->n) * g->d * exp(-g->a * r2);
  }
  return potential;
}

/* Outer radius where |c| r^l exp(-zeta r^2) drops to the cut-off, by
   Newton-Raphson on t + l ln r - zeta r^2; returns r^2; util.c:133-172.
   *ok = 0 where the reference's assert would fire. */
double orc_gto_radius(double c, double zeta, int l, double cutoff, int *ok) {
  const int N = 40;
  double delta, dg, guess, zr, r;
  const double t = log(fabs(c) / fabs(cutoff));
  int converged = 0;
  dg = t / zeta;
  r = (dg > cutoff) ? dg : cutoff;
  *ok = 1;
  if (0 != l) {
    if (0 < l) {
      dg = sqrt(0.5 * l / fabs(c));
      guess = (dg > cutoff) ? dg : cutoff;
      if (guess > r) r = 0.5 * (r + guess);
    }
    for (int i = 0; i < N; i++) {
      zr = zeta * r;
      guess = t + l * log(r) - zr * r;
      delta = guess / (l / r - 2 * zr);
      dg = r - delta;
      r = (dg > cutoff) ? dg : cutoff;
      if (fabs(delta) < cutoff) {
        r *= r;
        converged = 1;
        break;
      }
    }
    if (!converged) *ok = 0;
  }
  return r;
}

/* Radius of a contracted shell from its most diffuse primitive with a
   non-zero coefficient; util.c:175-192. */
double orc_shell_radius_(int K, int l, const double *d, const double *a, double zero, int *ok) {
  double zeta = a[0];
  double c = fabs(d[0]);
  for (int i = 1; i < K; i++) {
    if (a[i] < zeta && d[i] != 0.0) {
      zeta = a[i];
      c = fabs(d[i]);
    }
  }
  return sqrt(orc_gto_radius(c, zeta, l, zero, ok));
}

/* ---------------------------------------------------------------------- */
/* Exported helpers for the unit tests.                                   */

void orc_grid_ps93(int maxPoints, double *x, double *w, int *order) {
  orc_grid *g = orc_grid_new_ps93(maxPoints, 0.0);
  orc_map_kk(g);
  memcpy(x, g->x, sizeof(double) * (size_t)g->order);
  memcpy(w, g->w, sizeof(double) * (size_t)g->order);
  *order = g->order;
  orc_grid_free(g);
}

void orc_grid_psm92(int maxPoints, double *x, double *w, int *order) {
  orc_grid *g = orc_grid_new_psm92(maxPoints, 0.0);
  memcpy(x, g->x, sizeof(double) * (size_t)g->order);
  memcpy(w, g->w, sizeof(double) * (size_t)g->order);
  *order = g->order;
  orc_grid_free(g);
}

int orc_bessel_table(int lMax, int N, int cutoff, double accuracy, double *K) {
  orc_bessel b;
  int rc = orc_bessel_init(&b, lMax, N, cutoff, accuracy);
  if (rc == 0) memcpy(K, b.K, sizeof(double) * (size_t)(lMax + 1) * (N + 1));
  orc_bessel_release(&b);
  return rc;
}

void orc_bessel_eval(int lMax, double accuracy, int lmax, double z, double *K) {
  orc_bessel b;
  orc_bessel_init(&b, lMax, 1600, 200, accuracy);
  orc_bessel_weighted(&b, 1, lmax, z, K);
  orc_bessel_release(&b);
}

void orc_real_sph(int lmax, double theta, double phi, double *out) {
  double *fac = orc_fac_table(2 * lmax + 2, 1);
  double *dfac = orc_fac_table(2 * lmax + 2, 2);
  orc_rsh(lmax, theta, phi, fac, dfac, out);
  free(fac);
  free(dfac);
}

void orc_spherical_coordinates(const double *xyz, double *rtp) { orc_sph_coords(xyz, rtp); }

double orc_shell_radius(int K, int l, const double *d, const double *a, double zero) {
  int ok;
  return orc_shell_radius_(K, l, d, a, zero, &ok);
}

/* Quadrature hooks for the unit tests: f(r) integrated by the PS93
   sequence on the Krack-Koster grid, or by PSM92 on the FM06-mapped large
   grid (zeta_p, P), with the integrators' own convergence logic. */
typedef struct {
  double (*f)(double);
  const double *x;
} orc_hook;

static double hook_integrand(int index, const void *p) {
  const orc_hook *h = p;
  return h->f(h->x[index]);
}

double orc_quad_ps93(double (*f)(double), double tol, int *rc) {
  orc_grid *g = orc_grid_new_ps93(128, tol);
  orc_map_kk(g);
  orc_hook h = {f, g->x};
  *rc = orc_integrate_ps93(hook_integrand, &h, g);
  double I = g->I;
  orc_grid_free(g);
  return I;
}

double orc_quad_psm92(int maxPoints, double zeta_p, double P, double (*f)(double), double tol, int *rc) {
  orc_grid *g = orc_grid_new_psm92(maxPoints, tol);
  orc_map_fm06(g, zeta_p, P);
  orc_hook h = {f, g->x};
  *rc = orc_integrate_psm92(hook_integrand, &h, g);
  double I = g->I;
  orc_grid_free(g);
  return I;
}
