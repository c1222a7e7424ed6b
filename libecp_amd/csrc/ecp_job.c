/* Host side of a calculation: static tables, screening and the triplet list.
 *
 * A job is a set of molecules flattened into global atom / shell / primitive
 * / ECP-centre arrays.  For every molecule the host computes, exactly as the
 * reference does per ECP centre C (libecp.c:287-303):
 *   - the progressive cut-off of the semi-local potentials on the small grid
 *     (tabECP_FM06 with potentialScreening, type2.c:184-219, util.c:198-210);
 *   - each shell's radius and screened window (shellRadius / gtoRadius,
 *     util.c:133-192; basisSetScreening, type2.c:148-180);
 *   - the evaluated triplets (C, shell a, shell b >= a) in the reference's
 *     loop order C -> A -> B >= A -> sa -> sb with its skip rules and the
 *     pair window [max start, max end] (libecp.c:309-375).
 * Everything numerical per grid point runs on the device.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ecp_host.h"

static void set_err(char *err, size_t n, const char *msg) {
  if (err && n) snprintf(err, n, "%s", msg);
}

/* U_l(r) = sum over the channel's Gaussians of d r^n exp(-a r^2) (ecp.c:41-60). */
double ecp_eval_ecp_host(const ecp_gauss *g, int ng, int l, double r) {
  double potential = 0.0;
  const double r2 = r * r;
  for (int i = 0; i < ng; i++)
    if (l == g[i].l) potential += pow(r, g[i].n) * g[i].d * exp(-g[i].a * r2);
  return potential;
}

/* Newton-Raphson radius of the most diffuse primitive (util.c:133-192). */
double ecp_shell_radius(int K, int l, const double *d, const double *a, double zero, int *ok) {
  double zeta = a[0], c = fabs(d[0]);
  for (int i = 1; i < K; i++)
    if (a[i] < zeta && d[i] != 0.0) {
      zeta = a[i];
      c = fabs(d[i]);
    }
  const double t = log(fabs(c) / fabs(zero));
  double dg = t / zeta, r = (dg > zero) ? dg : zero, guess, delta, zr;
  *ok = 1;
  if (0 != l) {
    int converged = 0;
    if (0 < l) {
      dg = sqrt(0.5 * l / fabs(c));
      guess = (dg > zero) ? dg : zero;
      if (guess > r) r = 0.5 * (r + guess);
    }
    for (int i = 0; i < 40; i++) {
      zr = zeta * r;
      guess = t + l * log(r) - zr * r;
      delta = guess / (l / r - 2 * zr);
      dg = r - delta;
      r = (dg > zero) ? dg : zero;
      if (fabs(delta) < zero) {
        r *= r;
        converged = 1;
        break;
      }
    }
    if (!converged) *ok = 0;
  }
  return sqrt(r);
}

int ecp_mol_scan(const ecp_mol *m, ecp_mol_dims *d, char *err, size_t errlen) {
  memset(d, 0, sizeof(*d));
  if (m->nat <= 0 || !m->geom || !m->shellsBS || !m->shellsECP) {
    set_err(err, errlen, "empty molecule or missing arrays");
    return -1;
  }
  int ish = 0, iecp = 0;
  for (int i = 0; i < m->nat; i++) {
    if (m->shellsBS[i] < 0 || m->shellsECP[i] < 0) {
      set_err(err, errlen, "negative shell count");
      return -1;
    }
    for (int j = 0; j < m->shellsBS[i]; j++, ish++) {
      const int l = m->lBS[ish], K = m->KBS[ish];
      if (l < 0 || l > ECP_MAX_LBS || K <= 0) {
        set_err(err, errlen, "basis shell with unsupported l (0..5) or K <= 0");
        return -1;
      }
      if (l > d->maxLBS) d->maxLBS = l;
      d->nprim += K;
      d->nao += ECP_DIM_IJK(l);
    }
    if (m->shellsECP[i] > 0) {
      int L = 0;
      d->necp++;
      for (int j = 0; j < m->shellsECP[i]; j++, iecp++) {
        if (m->lECP[iecp] < 0 || m->KECP[iecp] < 0) {
          set_err(err, errlen, "invalid ECP shell");
          return -1;
        }
        if (m->lECP[iecp] > L) L = m->lECP[iecp];
        d->ngauss += m->KECP[iecp];
      }
      if (L < 1 || L > ECP_MAX_L) {
        set_err(err, errlen, "ECP local channel label must be 1..7 (an ECP needs a semi-local channel)");
        return -1;
      }
      if (L > d->maxLECP) d->maxLECP = L;
    }
  }
  d->nsh = ish;
  return 0;
}

void ecp_static_free(ecp_static *st) {
  free(st->bessel);
  free(st->fac);
  free(st->dfac);
  free(st->ijk);
  free(st->ijkIndex);
  free(st->cart2sph);
  free(st->poly2sph);
  free(st->omega);
  memset(st, 0, sizeof(*st));
}

/* Tables for a handle whose largest ECP label is maxL and whose largest basis
   l (plus derivative order) is maxAlpha; libECP_init, libecp.c:129-198. */
int ecp_static_init(ecp_static *st, int maxL, int maxAlpha, const int *ordering, int lmaxOrdering,
                    int largeGridOrder, double tolerance, double accuracy, char *err, size_t errlen) {
  memset(st, 0, sizeof(*st));
  const int maxLambda = maxL - 1 + maxAlpha;
  const int tmDim = maxLambda + maxAlpha;
  st->tolerance = tolerance;
  st->accuracy = accuracy;
  st->lnAcc1 = log(accuracy) - 2;
  st->lnAcc2 = log(1.0E-14) - 2;
  /* at least 0..5: the reference's factorials(1) leaves 1! = 0 and its
     TM_DIM reads fac[3] when tmDim = 0 (util.c:65-70, transformations.h:13) */
  st->fac = ecp_factorials(2 * tmDim + 5, 1);
  st->dfac = ecp_factorials(2 * tmDim + 5, 2);
  if (!st->fac || !st->dfac) {
    set_err(err, errlen, "factorial table overflow");
    ecp_static_free(st);
    return -1;
  }
  if (!ordering) {
    st->ijk = ecp_cart_order(tmDim);
    st->ijkDim = tmDim + 1;
    st->ijkIndex = ecp_cart_index(tmDim, st->ijk);
  } else {
    if (lmaxOrdering < tmDim + 1) {
      set_err(err, errlen, "custom shell ordering does not cover lmax >= maxLambda + maxAlpha + 1");
      ecp_static_free(st);
      return -1;
    }
    st->ijk = malloc(sizeof(int) * 3 * (size_t)ECP_DIM_C(lmaxOrdering));
    if (st->ijk) memcpy(st->ijk, ordering, sizeof(int) * 3 * (size_t)ECP_DIM_C(lmaxOrdering));
    st->ijkDim = lmaxOrdering + 1;
    st->ijkIndex = st->ijk ? ecp_cart_index(lmaxOrdering, st->ijk) : NULL;
  }
  if (!st->ijk || !st->ijkIndex) {
    set_err(err, errlen, "out of host memory for the Cartesian ordering tables");
    ecp_static_free(st);
    return -1;
  }
  st->tmDim = tmDim;
  st->cart2sph = ecp_cart2sph(tmDim, st->ijk, st->fac);
  st->poly2sph = st->cart2sph ? ecp_poly2sph(st->cart2sph, tmDim, st->ijk, st->dfac) : NULL;
  st->omL = maxL;
  st->omLambda = maxLambda;
  st->omAlpha = maxAlpha;
  st->omega = st->poly2sph ? ecp_omega(maxL, maxLambda, maxAlpha, st->ijk, st->ijkIndex, st->ijkDim,
                                       st->cart2sph, st->poly2sph, tmDim, st->dfac)
                           : NULL;
  if (!st->omega) {
    set_err(err, errlen, "out of host memory for the angular tables");
    ecp_static_free(st);
    return -1;
  }
  if (ecp_grid_ps93_kk(st->xs, st->ws) != ECP_SMALL_ORDER) {
    set_err(err, errlen, "small grid construction failed");
    ecp_static_free(st);
    return -1;
  }
  st->lorder = ecp_grid_psm92(largeGridOrder, st->xl, st->wl);
  if (st->lorder < ECP_SMALL_ORDER) {
    /* the reference sizes type-1 arrays by the large order and indexes the
       383-point grid (type1.c:114-125): orders below 383 are invalid */
    set_err(err, errlen, "largeGridOrder must give a PSM92 grid of 383..2047 points (e.g. 1024)");
    ecp_static_free(st);
    return -1;
  }
  ecp_schedule_ps93(&st->s93, st->p93l, st->p93r);
  ecp_schedule_psm92(&st->s92, st->lorder, st->p92l, st->p92r);
  if (maxL + maxAlpha + 6 + 1 > ECP_BESSEL_ROWS) {
    set_err(err, errlen, "angular momenta exceed the Bessel table of this build");
    ecp_static_free(st);
    return -1;
  }
  st->bessel = malloc(sizeof(double) * (ECP_BESSEL_N + 1) * ECP_BESSEL_STRIDE);
  const int brc = st->bessel ? ecp_bessel_table(accuracy, ECP_BESSEL_N, 200, ECP_BESSEL_ROWS, st->bessel) : -1;
  if (brc) {
    set_err(err, errlen, brc < 0 ? "out of host memory for the Bessel table"
                                 : "Bessel function tabulation did not converge");
    ecp_static_free(st);
    return -1;
  }
  return 0;
}

void ecp_job_free(ecp_job *job) {
  free(job->atoms);
  free(job->atom_maxl);
  free(job->atom_shell0);
  free(job->atom_nsh);
  free(job->shells);
  free(job->pa);
  free(job->pd);
  free(job->cent);
  free(job->gauss);
  free(job->win);
  free(job->trip);
  free(job->trip_mol_begin);
  free(job->mol_atom0);
  memset(job, 0, sizeof(*job));
}

typedef struct {
  ecp_triplet *v;
  int n, cap;
} trip_vec;

static int trip_push(trip_vec *tv, const ecp_triplet *t) {
  if (tv->n == tv->cap) {
    int nc = tv->cap ? 2 * tv->cap : 1024;
    ecp_triplet *nv = realloc(tv->v, sizeof(ecp_triplet) * (size_t)nc);
    if (!nv) return -1;
    tv->v = nv;
    tv->cap = nc;
  }
  tv->v[tv->n++] = *t;
  return 0;
}

/* Derivative shifts of the basis angular momenta, in the reference's order
   (libecp.c:243-250): first derivatives (+1,0) (-1,0) (0,+1) (0,-1); second
   derivatives the ten pairs below. */
static const int kShiftN[3] = {1, 4, 10};
static const int kShiftA[3][10] = {{0}, {+1, -1, 0, 0}, {+2, -2, +1, +1, -1, -1, 0, 0, +1, 0}};
static const int kShiftB[3][10] = {{0}, {0, 0, +1, -1}, {0, 0, +1, -1, +1, -1, +2, -2, 0, +1}};

int ecp_job_build(const ecp_mol *mols, int nmol, int deriv, const ecp_static *st, ecp_job *job, char *err,
                  size_t errlen) {
  memset(job, 0, sizeof(*job));
  ecp_mol_dims *dims = calloc((size_t)nmol, sizeof(ecp_mol_dims));
  if (!dims) {
    snprintf(err, errlen, "out of host memory");
    return -1;
  }
  int nat = 0, nsh = 0, nprim = 0, ncent = 0, ngauss = 0, nwin = 0;
  int maxLBS = 0, maxL = 0;
  for (int m = 0; m < nmol; m++) {
    if (ecp_mol_scan(&mols[m], &dims[m], err, errlen)) {
      free(dims);
      return -1;
    }
    nat += mols[m].nat;
    nsh += dims[m].nsh;
    nprim += dims[m].nprim;
    ncent += dims[m].necp;
    ngauss += dims[m].ngauss;
    nwin += dims[m].necp * dims[m].nsh;
    if (dims[m].maxLBS > maxLBS) maxLBS = dims[m].maxLBS;
    if (dims[m].maxLECP > maxL) maxL = dims[m].maxLECP;
  }
  if (deriv < 0 || deriv > 2) {
    set_err(err, errlen, "derivative order must be 0, 1 or 2");
    free(dims);
    return -1;
  }
  if (maxLBS + deriv > ECP_MAX_LBS) {
    set_err(err, errlen, "basis angular momentum plus derivative order exceeds this build (at most 6)");
    free(dims);
    return -1;
  }
  if (maxL - 1 + maxLBS + deriv > st->omLambda || maxLBS + deriv > st->omAlpha || maxL > st->omL) {
    set_err(err, errlen, "molecule exceeds the angular dimensions of the static tables");
    free(dims);
    return -1;
  }
  job->nmol = nmol;
  job->deriv = deriv;
  job->nat = nat;
  job->nsh = nsh;
  job->nprim = nprim;
  job->ncent = ncent;
  job->ngauss = ngauss;
  job->nwin = nwin;
  job->maxL = maxL;
  job->FL = maxL + maxLBS + deriv;          /* rows lambda = 0 .. L-1+l_s+run */
  job->UK = 2 * (maxLBS + deriv) + 1;       /* r^k, k = 0 .. la+lb (shifted) */
  job->angL = maxL;
  job->ang_lam = maxL + maxLBS + deriv;     /* lambda = 0 .. L-1+lmax_a+n */
  job->ang_c = ECP_DIM_C(maxLBS + deriv);
  job->atoms = calloc((size_t)nat + 1, sizeof(ecp_xyz));
  job->atom_maxl = calloc((size_t)nat + 1, sizeof(int));
  job->atom_shell0 = calloc((size_t)nat + 1, sizeof(int));
  job->atom_nsh = calloc((size_t)nat + 1, sizeof(int));
  job->shells = calloc((size_t)nsh + 1, sizeof(ecp_shell));
  job->pa = calloc((size_t)nprim + 1, sizeof(double));
  job->pd = calloc((size_t)nprim + 1, sizeof(double));
  job->cent = calloc((size_t)ncent + 1, sizeof(ecp_centre));
  job->gauss = calloc((size_t)ngauss + 1, sizeof(ecp_gauss));
  job->win = calloc((size_t)nwin + 1, sizeof(ecp_window));
  job->trip_mol_begin = calloc((size_t)nmol + 1, sizeof(int));
  job->mol_atom0 = calloc((size_t)nmol + 1, sizeof(int));
  double *radius = calloc((size_t)nsh + 1, sizeof(double));
  if (!job->atoms || !job->atom_maxl || !job->atom_shell0 || !job->atom_nsh || !job->shells || !job->pa ||
      !job->pd || !job->cent || !job->gauss || !job->win || !job->trip_mol_begin || !job->mol_atom0 || !radius) {
    snprintf(err, errlen, "out of host memory");
    free(radius);
    free(dims);
    ecp_job_free(job);
    return -1;
  }
  trip_vec tv = {0};
  int a0 = 0, s0 = 0, p0 = 0, c0 = 0, g0 = 0, w0 = 0;
  size_t blk = 0, ul = 0, ut = 0, fo = 0, ango = 0;
  int maxK = 0, rc = 0;
  for (int m = 0; m < nmol && !rc; m++) {
    const ecp_mol *M = &mols[m];
    /* atoms, shells, primitives */
    int sh = s0, pr = p0;
    for (int i = 0; i < M->nat; i++) {
      const int A = a0 + i;
      job->atoms[A].x = M->geom[3 * i];
      job->atoms[A].y = M->geom[3 * i + 1];
      job->atoms[A].z = M->geom[3 * i + 2];
      job->atom_shell0[A] = sh;
      job->atom_nsh[A] = M->shellsBS[i];
      for (int j = 0; j < M->shellsBS[i]; j++, sh++) {
        ecp_shell *S = &job->shells[sh];
        S->atom = A;
        S->l = M->lBS[sh - s0];
        S->K = M->KBS[sh - s0];
        S->prim = pr;
        S->mol = m;
        S->local = sh - s0;
        if (S->l > job->atom_maxl[A]) job->atom_maxl[A] = S->l;
        if (S->K > maxK) maxK = S->K;
        for (int k = 0; k < S->K; k++, pr++) {
          job->pa[pr] = M->aBS[pr - p0];
          job->pd[pr] = M->dBS[pr - p0];
        }
        int ok = 1;
        radius[sh] = ecp_shell_radius(S->K, S->l, job->pd + S->prim, job->pa + S->prim, 1.0E-14, &ok);
        if (!ok) {
          set_err(err, errlen, "no outer radius found for a basis shell (Newton iteration did not converge)");
          rc = -1;
        }
      }
    }
    /* ECP centres and Gaussians, unpacked as libecp.c:92-127 */
    int ie = 0, in = 0, cc = c0, gg = g0;
    for (int i = 0; i < M->nat; i++) {
      if (M->shellsECP[i] <= 0) continue;
      ecp_centre *Cn = &job->cent[cc];
      Cn->atom = a0 + i;
      Cn->mol = m;
      Cn->g0 = gg;
      Cn->L = 0;
      for (int j = 0; j < M->shellsECP[i]; j++, ie++) {
        if (M->lECP[ie] > Cn->L) Cn->L = M->lECP[ie];
        for (int k = 0; k < M->KECP[ie]; k++, in++, gg++) {
          job->gauss[gg].l = M->lECP[ie];
          job->gauss[gg].n = M->nECP[in];
          job->gauss[gg].d = M->dECP[in];
          job->gauss[gg].a = M->aECP[in];
        }
      }
      Cn->ng = gg - Cn->g0;
      cc++;
    }
    job->trip_mol_begin[m] = tv.n;
    job->mol_atom0[m] = a0;
    /* per centre: potential cut-offs, windows, triplets */
    for (int c = c0; c < cc && !rc; c++) {
      ecp_centre *Cn = &job->cent[c];
      const ecp_xyz rC = job->atoms[Cn->atom];
      int end = ECP_SMALL_ORDER - 1;
      for (int l = 0; l <= ECP_MAX_L; l++) Cn->end[l] = -1;
      for (int l = 0; l < Cn->L; l++) {
        int e = -1;
        for (int i = end; i >= 0; i--)
          if (fabs(ecp_eval_ecp_host(job->gauss + Cn->g0, Cn->ng, l, st->xs[i])) > 1.0E-14) {
            e = i;
            break;
          }
        end = e;
        Cn->end[l] = end;
      }
      Cn->UL_off = (int)ul;
      ul += ECP_SMALL_ORDER;
      Cn->Ut_off = (int)ut;
      ut += (size_t)maxL * job->UK * ECP_SMALL_ORDER;
      Cn->F_off = (int)fo;
      fo += (size_t)(deriv + 1) * dims[m].nsh * job->FL * ECP_SMALL_ORDER;
      Cn->win_off = w0 + (c - c0) * dims[m].nsh;
      Cn->atom0 = a0;
      Cn->nat = M->nat;
      Cn->shell0 = s0;
      Cn->nsh = dims[m].nsh;
      Cn->ang_off = (int)ango;
      ango += (size_t)M->nat * job->ang_lam * job->angL * job->angL * job->ang_c;
      /* basis screening of each shell on [0, end] (type2.c:148-180) */
      for (int s = 0; s < dims[m].nsh; s++) {
        const ecp_shell *S = &job->shells[s0 + s];
        const ecp_xyz rA = job->atoms[S->atom];
        const double x = rC.x - rA.x, y = rC.y - rA.y, z = rC.z - rA.z;
        const double dAC = sqrt(x * x + y * y + z * z);
        const double rmin = dAC - radius[s0 + s], rmax = dAC + radius[s0 + s];
        ecp_window *W = &job->win[Cn->win_off + s];
        int j;
        for (j = end; j >= 0; j--)
          if (st->xs[j] < rmin) break;
        W->start = j + 1;
        for (j = end; j >= 0; j--)
          if (st->xs[j] <= rmax) break;
        W->end = (j < 0 || st->xs[j] > rmax) ? -1 : j;
        W->skip = !(W->end >= W->start);
      }
      /* triplets in the reference order */
      for (int A = 0; A < M->nat && !rc; A++) {
        const int gA = a0 + A;
        for (int B = A; B < M->nat && !rc; B++) {
          const int gB = a0 + B;
          for (int s1 = 0; s1 < job->atom_nsh[gA]; s1++) {
            const int sa = job->atom_shell0[gA] + s1;
            const ecp_window *Wa = &job->win[Cn->win_off + (sa - s0)];
            if (Wa->skip) continue;
            for (int s2 = (A == B) ? s1 : 0; s2 < job->atom_nsh[gB]; s2++) {
              const int sb = job->atom_shell0[gB] + s2;
              const ecp_window *Wb = &job->win[Cn->win_off + (sb - s0)];
              if (Wb->skip) continue;
              ecp_triplet t;
              memset(&t, 0, sizeof(t));
              t.centre = c;
              t.sa = sa;
              t.sb = sb;
              t.gs = Wa->start > Wb->start ? Wa->start : Wb->start;
              t.ge = Wa->end > Wb->end ? Wa->end : Wb->end;
              if (!(t.gs < t.ge)) continue;
              /* derivatives: A = B = C contributes nothing (libecp.c:305-306) */
              if (deriv > 0 && gA == Cn->atom && gB == Cn->atom) continue;
              const int la0 = job->shells[sa].l, lb0 = job->shells[sb].l;
              /* one block pair per derivative shift (libecp.c:322-330) */
              for (int k = 0; k < kShiftN[deriv] && !rc; k++) {
                const int sha = kShiftA[deriv][k], shb = kShiftB[deriv][k];
                if (la0 < -sha || lb0 < -shb || (deriv == 1 && Cn->atom == gA && sha != 0) ||
                    (deriv == 1 && Cn->atom == gB && shb != 0))
                  continue;
                t.shifta = sha;
                t.shiftb = shb;
                t.la = la0 + sha;
                t.lb = lb0 + shb;
                t.powa = sha > 0 ? sha : 0;
                t.powb = shb > 0 ? shb : 0;
                /* second derivatives: the (+1, 0) and (0, +1) blocks are
                   shifted back on the unshifted side (libecp.c:364-369) */
                t.oa = (deriv == 2 && sha == 1 && shb == 0) ? la0 : t.la;
                t.ob = (deriv == 2 && sha == 0 && shb == 1) ? lb0 : t.lb;
                const int nb = ECP_DIM_IJK(t.oa) * ECP_DIM_IJK(t.ob);
                t.blk1 = (int)blk;
                t.blk2 = (int)(blk + nb);
                blk += 2 * (size_t)nb;
                if (trip_push(&tv, &t)) {
                  set_err(err, errlen, "out of host memory");
                  rc = -1;
                }
              }
            }
          }
        }
      }
    }
    a0 += M->nat;
    s0 = sh;
    p0 = pr;
    c0 = cc;
    g0 = gg;
    w0 += dims[m].necp * dims[m].nsh;
  }
  job->trip_mol_begin[nmol] = tv.n;
  job->mol_atom0[nmol] = a0;
  job->trip = tv.v;
  job->ntrip = tv.n;
  job->nblk = blk;
  job->maxK = maxK;
  job->nUL = ul;
  job->nUt = ut;
  job->nF = fo;
  job->nang = ango;
  free(radius);
  free(dims);
  if (rc) ecp_job_free(job);
  return rc;
}

/* Deep copy of a molecule's input arrays into one allocation (released by
   ecp_mol_release); validates it first.  Fills *d with its dimensions. */
int ecp_mol_copy(const ecp_mol *src, ecp_mol *dst, ecp_mol_dims *d, char *err, size_t errlen) {
  memset(dst, 0, sizeof(*dst));
  if (ecp_mol_scan(src, d, err, errlen)) return -1;
  int necps = 0;
  for (int i = 0; i < src->nat; i++) necps += src->shellsECP[i];
  const size_t nd = 3 * (size_t)src->nat + 3 * (size_t)d->ngauss + 2 * (size_t)d->nprim;
  const size_t ni = 2 * (size_t)src->nat + 2 * (size_t)necps + 2 * (size_t)d->nsh;
  double *fb = malloc(sizeof(double) * nd + sizeof(int) * ni + 8);
  if (!fb) {
    set_err(err, errlen, "out of host memory");
    return -1;
  }
  double *f = fb;
  int *ib = (int *)(fb + nd), *ip = ib;
#define TAKE_D(dstp, s, n)                        \
  do {                                            \
    memcpy(f, (s), sizeof(double) * (size_t)(n)); \
    (dstp) = f;                                   \
    f += (n);                                     \
  } while (0)
#define TAKE_I(dstp, s, n)                         \
  do {                                             \
    memcpy(ip, (s), sizeof(int) * (size_t)(n));    \
    (dstp) = ip;                                   \
    ip += (n);                                     \
  } while (0)
  dst->nat = src->nat;
  TAKE_D(dst->geom, src->geom, 3 * (size_t)src->nat);
  TAKE_D(dst->nECP, src->nECP, d->ngauss);
  TAKE_D(dst->dECP, src->dECP, d->ngauss);
  TAKE_D(dst->aECP, src->aECP, d->ngauss);
  TAKE_D(dst->dBS, src->dBS, d->nprim);
  TAKE_D(dst->aBS, src->aBS, d->nprim);
  TAKE_I(dst->shellsECP, src->shellsECP, src->nat);
  TAKE_I(dst->lECP, src->lECP, necps);
  TAKE_I(dst->KECP, src->KECP, necps);
  TAKE_I(dst->shellsBS, src->shellsBS, src->nat);
  TAKE_I(dst->lBS, src->lBS, d->nsh);
  TAKE_I(dst->KBS, src->KBS, d->nsh);
#undef TAKE_D
#undef TAKE_I
  return 0;
}

void ecp_mol_release(ecp_mol *m) {
  free((void *)m->geom);
  memset(m, 0, sizeof(*m));
}

/* AO offset of every shell of a job inside its own molecule's matrix:
   atom-major, IJK(l) functions per shell (getIntegrals.c:59-72). */
void ecp_job_ao_offsets(const ecp_job *J, int *aoff) {
  int cur = -1, off = 0;
  for (int s = 0; s < J->nsh; s++) {
    if (J->shells[s].mol != cur) {
      cur = J->shells[s].mol;
      off = 0;
    }
    aoff[s] = off;
    off += ECP_DIM_IJK(J->shells[s].l);
  }
}

/* Adds molecule m's blocks into the upper triangle (row <= column) of its
   AO matrix, as getIntegrals' callback does (getIntegrals.c:22-43): dense
   with row stride rowdim when packed_n == 0, else packed by rows (row i
   holds columns i..packed_n-1).  Stops at the first failed integration in
   callback order and returns its code (1 type 1, 2 type 2), else 0. */
int ecp_job_scatter(const ecp_job *J, int m, const int *aoff, const double *blocks, const unsigned char *status,
                    int rowdim, int packed_n, double *I) {
  for (int t = J->trip_mol_begin[m]; t < J->trip_mol_begin[m + 1]; t++) {
    const ecp_triplet *tr = &J->trip[t];
    const int la = J->shells[tr->sa].l, lb = J->shells[tr->sb].l;
    const int a0 = aoff[tr->sa], b0 = aoff[tr->sb];
    const int na = ECP_DIM_IJK(la), nb = ECP_DIM_IJK(lb);
    for (int type = 0; type < 2; type++) {
      if (status[2 * t + type]) return type + 1;
      const double *blk = blocks + (type ? tr->blk2 : tr->blk1);
      for (int i = 0; i < na; i++) {
        const size_t r = (size_t)(a0 + i);
        double *row = packed_n ? I + r * packed_n - r * (r - 1) / 2 - r : I + r * rowdim;
        for (int j = 0; j < nb; j++)
          if (a0 + i <= b0 + j) row[b0 + j] += blk[i * nb + j];
      }
    }
  }
  return 0;
}
