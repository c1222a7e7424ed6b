// HIP device layer of the ECP integral engine, written for gfx950 (MI355X).
//
// Pipeline of one job (all FP64, one HIP stream):
//   k_utab   U_L(r_n) and r_n^k U_l(r_n) on the 383-point small grid, per centre
//            (evalECP / tabECP_FM06: ecp.c:41-60, type2.c:184-219)
//   k_ftab   F_lambda(r_n) per (centre, shell) on its screened window
//            (calcF_FM06: type2.c:222-310)
//   k_ang    per-(centre, atom) angular tables (evalAngularIntegrals:
//            angular_integrals.c:104-142)
//   k_type1  type-1 radial integrals Q[l1][l2] per (triplet, primitive pair)
//            with the PS93 sequence (calcQ: type1.c:94-203)
//   k_type2  type-2 radial integrals T_l[l1][l2][l3] per entry with PS93
//            (calcT_fast_FM06: type2.c:336-381)
//   k_fb1 / k_fb2  PSM92 large-grid redo of failed entries on FM06-mapped
//            grids (type1.c:205-252, calcT_FM06: type2.c:417-528)
//   k_assemble  type-1 and type-2 blocks, binomial shift and norms
//            (calcChi: type1.c:211-302, calcGamma: type2.c:532-631,
//             calcPolynomials: util.c:246-334)
//
// The adaptive quadratures are evaluated lazily: one lane owns one integral
// (or one primitive pair's set of integrals) and visits the grid points in
// the integrator's own level order, replaying its recurrences and its
// convergence test after every level.  A lane stops at convergence, so only
// the points the reference's result depends on are evaluated; the reference
// tabulates every window point up front.  Summation order inside a level and
// the recurrences are those of gc_integrators.c:156-217 (PS93) and :38-86
// (PSM92).
//
// Large-grid fallback semantics: points past a screening break contribute
// zero (the reference reuses stale buffers there, type2.c:443-481).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ecp_host.h"

#define ECP_CHECK(call)                                                            \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      snprintf(err, errlen, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
               __FILE__, __LINE__);                                                \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

namespace {

constexpr int SO = ECP_SMALL_ORDER;
constexpr int kFbChunk = 8;       // failed type-2 entries per fallback lane
constexpr int kMaxLMB = ECP_MAX_L - 1 + ECP_MAX_LBS;  // largest Bessel order of F / K_a / K_b

struct DevTables {
  const double *xs, *ws, *xl, *wl;
  const int *p93l, *p93r, *p92l, *p92r;
  const double *bessel;
  const double *poly2sph;
  const double *omega;
  const int *ijk, *ijkIndex;
  const double *binom;  // binom[n*16+k], n,k <= 15
  int ijkDim, tmDim, lorder;
  int omL, omLambda, omAlpha;
  double tol, lnAcc1, lnAcc2, accuracy;
  ecp_schedule s93, s92;
};

struct DevJob {
  const ecp_xyz *atoms;
  const int *atom_maxl;
  const ecp_shell *shells;
  const double *pa, *pd;
  const ecp_centre *cent;
  const ecp_gauss *gauss;
  const ecp_window *win;
  const ecp_triplet *trip;
  int ntrip;
  double *UL, *Ut, *F, *ang;
  int FL, UK, angL, ang_lam, ang_c;
  int maxPairs;
};

// ---------------------------------------------------------------------------
// numerical primitives

__device__ __forceinline__ double ecp_pow(double r, double n) {
  // exact special cases of a correctly rounded pow
  if (n == 2.0) return r * r;
  if (n == 1.0) return r;
  if (n == 0.0) return 1.0;
  return pow(r, n);
}

__device__ double eval_ecp(const ecp_gauss *__restrict__ g, int ng, int l, double r) {
  double pot = 0.0;
  const double r2 = r * r;
  for (int i = 0; i < ng; i++) {
    if (g[i].l == l) pot += ecp_pow(r, g[i].n) * g[i].d * exp(-g[i].a * r2);
  }
  return pot;
}

template <int J>
struct RecC {
  static constexpr double v = double(J) / (2.0 * J + 1.0);
};

// K_0..K_LM at z (bessel.c:102-199): leading term below 1e-7, 5-term Taylor
// series about the nearest table abscissa below 16 (derivatives by the
// recurrence K_j' = j/(2j+1)(K_{j-1} - K_{j+1}) - K_j + K_{j+1}), asymptotic
// series above.
template <int LM>
__device__ __forceinline__ void bessel_k(const double *__restrict__ tab, double z, double (&K)[LM + 1]) {
  if (z < 1.0E-7) {
    if (z <= 0) {
      K[0] = 1.0;
#pragma unroll
      for (int l = 1; l <= LM; l++) K[l] = 0.0;
    } else {
      K[0] = 1 - z;
#pragma unroll
      for (int l = 1; l <= LM; l++) K[l] = K[l - 1] * z / (2 * l + 1);
    }
  } else if (z < 16.0) {
    constexpr int ML = LM + 5;
    const int index = (int)floor(z * 100.0 + 0.5);
    const double dz = z - index / 100.0;
    const double *row = tab + (size_t)index * ECP_BESSEL_STRIDE;
    double d[ML + 2];
#pragma unroll
    for (int l = 0; l <= ML; l++) d[l] = row[l];
    d[ML + 1] = 0.0;
#pragma unroll
    for (int l = 0; l <= LM; l++) K[l] = d[l];
    double scale = 1.0;
#pragma unroll
    for (int i = 1; i <= 5; i++) {
      const int top = ML - i;
      double prev = d[0];
      d[0] = d[1] - d[0];
#pragma unroll
      for (int j = 1; j <= top; j++) {
        const double cur = d[j];
        double cj = double(j) / (2.0 * j + 1.0);
        d[j] = cj * (prev - d[j + 1]) - cur + d[j + 1];
        prev = cur;
      }
      scale = scale * dz / i;
#pragma unroll
      for (int j = 0; j <= LM; j++) K[j] += scale * d[j];
    }
  } else {
    double A[LM + 1];
    A[0] = 0.5 / z;
#pragma unroll
    for (int l = 0; l <= LM; l++) K[l] = A[0];
#pragma unroll
    for (int l = 1; l <= LM; l++) {
      double f = l * (l + 1);
#pragma unroll
      for (int i = 1; i < l; i++) {
        K[l] += f * A[i];
        f *= (l + i + 1) * (l - i);
      }
      A[l] = -A[0] * A[l - 1] / l;
      K[l] += f * A[l];
    }
  }
}

// Real spherical harmonics without Condon-Shortley phase
// (spherical_harmonics.c:15-114); out[l*l + l + m], m = -l..l.
__device__ void real_sph(int lmax, double theta, double phi, const double *__restrict__ fac,
                         const double *__restrict__ dfac, double *out, double *P, double *s, double *c) {
  const double x = cos(theta);
  for (int i = 0; i < (lmax + 1) * (lmax + 1); i++) P[i] = 0.0;
  for (int i = 0; i < lmax + 2; i++) s[i] = c[i] = 0.0;
#define RSH(l, m) ((l) * (l) + (l) + (m))
  if (1.0 == x) {
    for (int l = 0; l <= lmax; l++) P[RSH(l, 0)] = 1.0;
  } else if (-1.0 == x) {
    P[0] = 1.0;
    for (int l = 1; l <= lmax; l++) P[RSH(l, 0)] = -P[RSH(l - 1, 0)];
  } else {
    s[1] = sqrt(1.0 - x * x);
    for (int l = 2; l <= lmax; l++) s[l] = s[l - 1] * s[1];
    for (int l = 0; l <= lmax; l++) {
      if (l == 0) {
        P[RSH(0, 0)] = 1.0;
        continue;
      }
      P[RSH(l, l)] = s[l] * dfac[2 * l - 1];
      int m = l - 1;
      P[RSH(l, m)] = x * (2 * m + 1) * P[RSH(l - 1, m)];
      if (l > 1)
        for (m = 0; m <= l - 2; m++)
          P[RSH(l, m)] = (x * (2 * l - 1) * P[RSH(l - 1, m)] - (l + m - 1) * P[RSH(l - 2, m)]) / (l - m);
    }
  }
  for (int l = 0; l <= lmax; l++) {
    const double norm0 = sqrt((2.0 * l + 1.0) / (2.0 * M_PI));
    P[RSH(l, 0)] = norm0 * P[RSH(l, 0)];
    for (int m = 1; m <= l; m++) P[RSH(l, m)] = (sqrt(fac[l - m] / fac[l + m]) * norm0) * P[RSH(l, m)];
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
    out[RSH(l, 0)] = P[RSH(l, 0)] / sqrt(2.0);
    for (int m = 1; m <= l; m++) {
      out[RSH(l, -m)] = P[RSH(l, m)] * s[m];
      out[RSH(l, m)] = P[RSH(l, m)] * c[m];
    }
  }
#undef RSH
}

// (r, theta, phi) of util.c:74-106
__device__ void sph_coords(double x, double y, double z, double &r, double &theta, double &phi) {
  const double acc = 1.0E-14;
  r = sqrt(x * x + y * y + z * z);
  theta = (r < acc) ? 0.0 : acos(z / r);
  if (fabs(x) < acc) {
    if (fabs(y) < acc) phi = 0.0;
    else if (y < 0.0) phi = 1.5 * M_PI;
    else phi = 0.5 * M_PI;
  } else {
    phi = (x > 0.0) ? atan(y / x) : atan(y / x) + M_PI;
  }
}

__device__ __forceinline__ double dist(const ecp_xyz &a, const ecp_xyz &b) {
  const double x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
  return sqrt(x * x + y * y + z * z);
}

// ---------------------------------------------------------------------------
// per-centre potential tables

__global__ __launch_bounds__(384) void k_utab(DevTables T, DevJob J) {
  const int c = blockIdx.x, n = threadIdx.x;
  if (n >= SO) return;
  const ecp_centre C = J.cent[c];
  const ecp_gauss *g = J.gauss + C.g0;
  const double r = T.xs[n];
  J.UL[C.UL_off + n] = eval_ecp(g, C.ng, C.L, r);
  for (int l = 0; l < C.L; l++) {
    double P = (n <= C.end[l]) ? eval_ecp(g, C.ng, l, r) : 0.0;
    double *dst = J.Ut + C.Ut_off + (size_t)l * J.UK * SO + n;
    dst[0] = P;
    for (int k = 1; k < J.UK; k++) {
      P = P * r;
      dst[(size_t)k * SO] = P;
    }
  }
}

// F[centre][shell][lambda][n] (type2.c:263-303), one block per non-skipped
// (centre, shell); the whole slab is written (zeros outside the window).
template <int LM>
__device__ void ftab_slab(const DevTables &T, const DevJob &J, const ecp_centre &C, int sl, const ecp_window &W) {
  const ecp_shell S = J.shells[C.shell0 + sl];
  const double dAC = dist(J.atoms[C.atom], J.atoms[S.atom]);
  double *slab = J.F + C.F_off + (size_t)sl * J.FL * SO;
  for (int n = threadIdx.x; n < SO; n += blockDim.x) {
    double acc[LM + 1];
#pragma unroll
    for (int i = 0; i <= LM; i++) acc[i] = 0.0;
    if (n >= W.start && n < W.end) {
      const double r = T.xs[n];
      for (int p = 0; p < S.K; p++) {
        const double zeta = J.pa[S.prim + p], da = J.pd[S.prim + p];
        double K[LM + 1];
        bessel_k<LM>(T.bessel, 2.0 * zeta * dAC * r, K);
        double e = dAC - r;
        e = exp(-zeta * e * e);
#pragma unroll
        for (int i = 0; i <= LM; i++) acc[i] += da * K[i] * e;
      }
    }
#pragma unroll
    for (int i = 0; i <= LM; i++) slab[(size_t)i * SO + n] = acc[i];
    for (int i = LM + 1; i < J.FL; i++) slab[(size_t)i * SO + n] = 0.0;
  }
}

__global__ __launch_bounds__(128) void k_ftab(DevTables T, DevJob J, const int2 *__restrict__ items) {
  const int2 it = items[blockIdx.x];
  const ecp_centre C = J.cent[it.x];
  const ecp_window W = J.win[C.win_off + it.y];
  const int lm = C.L - 1 + J.shells[C.shell0 + it.y].l;
  switch (lm) {
#define FT(k) \
  case k:     \
    ftab_slab<k>(T, J, C, it.y, W); \
    break;
    FT(0) FT(1) FT(2) FT(3) FT(4) FT(5) FT(6) FT(7) FT(8) FT(9) FT(10) FT(11)
#undef FT
    default:
      break;
  }
}

// per-(centre, atom) angular table ang[lambda][LM(l,m)][C(la,c)]
__global__ __launch_bounds__(128) void k_ang(DevTables T, DevJob J, const int2 *__restrict__ items,
                                             const double *__restrict__ fac, const double *__restrict__ dfac) {
  __shared__ double rsph[(2 * ECP_MAX_L + 1) * (2 * ECP_MAX_L + 1)];
  __shared__ double Pw[(2 * ECP_MAX_L + 1) * (2 * ECP_MAX_L + 1)];
  __shared__ double sw[2 * ECP_MAX_L + 3], cw[2 * ECP_MAX_L + 3];
  const int2 it = items[blockIdx.x];
  const ecp_centre C = J.cent[it.x];
  const int A = C.atom0 + it.y;
  const int lmax_a = J.atom_maxl[A];
  const int lmax = C.L - 1 + lmax_a;
  if (threadIdx.x == 0) {
    const ecp_xyz a = J.atoms[A], c = J.atoms[C.atom];
    double r, th, ph;
    sph_coords(a.x - c.x, a.y - c.y, a.z - c.z, r, th, ph);
    real_sph(lmax, th, ph, fac, dfac, rsph, Pw, sw, cw);
  }
  __syncthreads();
  const int d1 = ECP_DIM_L(T.omLambda), d2 = ECP_DIM_C(T.omAlpha);
  const int nlm = C.L * C.L, nc = ECP_DIM_C(lmax_a);
  double *out = J.ang + C.ang_off + (size_t)it.y * J.ang_lam * J.angL * J.angL * J.ang_c;
  const int total = (lmax + 1) * nlm * nc;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int cc = idx % nc;
    const int lm = (idx / nc) % nlm;
    const int lambda = idx / (nc * nlm);
    double v = 0.0;
    for (int mu = 0; mu < 2 * lambda + 1; mu++)
      v += rsph[ECP_IDX_LM(lambda, mu)] * T.omega[((size_t)lm * d1 + ECP_IDX_LM(lambda, mu)) * d2 + cc];
    out[((size_t)lambda * J.angL * J.angL + lm) * J.ang_c + cc] = v;
  }
}

// ---------------------------------------------------------------------------
// type 1: PS93 over the window for all Q[l1][l2] of one primitive pair

__host__ __device__ constexpr int nq_of(int lab) {
  int n = 0;
  for (int l1 = 0; l1 <= lab; l1++) n += l1 / 2 + 1;
  return n;
}

struct T1Unit {
  int trip, pa, pb, qoff;
};

struct T1Geom {
  double za, zb, ca, cb, s, z, zd2, dAC, dBC;
  double P[3];
};

__device__ __forceinline__ void t1_geom(const DevJob &J, const ecp_triplet &tr, int pa, int pb, T1Geom &G) {
  const ecp_centre C = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const ecp_xyz rc = J.atoms[C.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const double rAC[3] = {ra.x - rc.x, ra.y - rc.y, ra.z - rc.z};
  const double rBC[3] = {rb.x - rc.x, rb.y - rc.y, rb.z - rc.z};
  G.dAC = dist(rc, ra);
  G.dBC = dist(rc, rb);
  G.za = J.pa[Sa.prim + pa];
  G.ca = J.pd[Sa.prim + pa];
  G.zb = J.pa[Sb.prim + pb];
  G.cb = J.pd[Sb.prim + pb];
  for (int p = 0; p < 3; p++) G.P[p] = 2.0 * (G.za * rAC[p] + G.zb * rBC[p]);
  G.s = sqrt(G.P[0] * G.P[0] + G.P[1] * G.P[1] + G.P[2] * G.P[2]);
  G.zd2 = -G.za * G.dAC * G.dAC - G.zb * G.dBC * G.dBC;
  G.z = -G.za - G.zb;
}

template <int LAB>
__global__ __launch_bounds__(64) void k_type1(DevTables T, DevJob J, const T1Unit *__restrict__ units, int nunits,
                                              double *__restrict__ qout, unsigned long long *__restrict__ qfail,
                                              int *__restrict__ fbq, int *__restrict__ fbcount) {
  constexpr int NQ = nq_of(LAB);
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  const T1Unit U = units[u];
  const ecp_triplet tr = J.trip[U.trip];
  const ecp_centre Cn = J.cent[tr.centre];
  T1Geom G;
  t1_geom(J, tr, U.pa, U.pb, G);
  const double C0 = G.ca * G.cb * exp(G.zd2);
  const double *__restrict__ UL = J.UL + Cn.UL_off;
  const int gs = tr.gs, ge = tr.ge;
  const double minExp = T.lnAcc1;

  double p[NQ], q[NQ], I[NQ], Tt[NQ];
  // w_n f_e(n) accumulated into acc[] (type1.c:78-88 integrand, tabulated only
  // on [gs, ge) as type1.c:121-130)
  auto accum = [&](int n, double (&acc)[NQ]) {
    if (n < gs || n >= ge) return;
    const double r = T.xs[n];
    const double expo = (G.z * r + G.s) * r;
    if (!(expo >= minExp)) return;
    const double g = exp(expo);
    double K[LAB + 1];
    bessel_k<LAB>(T.bessel, G.s * r, K);
    const double Un = UL[n], wn = T.ws[n];
    double rn = 1.0;
    int e = 0;
#pragma unroll
    for (int l1 = 0; l1 <= LAB; l1++) {
      const double a = (C0 * rn) * Un;
#pragma unroll
      for (int l2 = l1; l2 >= 0; l2 -= 2, e++) acc[e] += wn * ((a * K[l2]) * g);
      rn *= r;
    }
  };
#pragma unroll
  for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
  accum(T.s93.centre, Tt);
#pragma unroll
  for (int e = 0; e < NQ; e++) {
    p[e] = Tt[e];
    Tt[e] = 0.0;
  }
  accum(T.s93.first_l, Tt);
  {
    double t2[NQ];
#pragma unroll
    for (int e = 0; e < NQ; e++) t2[e] = 0.0;
    accum(T.s93.first_r, t2);
#pragma unroll
    for (int e = 0; e < NQ; e++) {
      q[e] = Tt[e] + t2[e];
      I[e] = p[e] + q[e];
    }
  }
  unsigned long long done = 0ull;
  const unsigned long long all = (NQ >= 64) ? ~0ull : ((1ull << NQ) - 1ull);
  for (int lv = 2; lv < T.s93.nlevels; lv++) {
    const int j = T.s93.lvl_j[lv], nn = T.s93.lvl_n[lv];
    int cnt = 0;
    for (int k = T.s93.lvl_begin[lv]; k < T.s93.lvl_begin[lv + 1]; k++) {
      const int li = T.p93l[k], ri = T.p93r[k];
#pragma unroll
      for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
      if (li >= gs) {
        cnt++;
        accum(li, Tt);
      }
      if (ri <= ge) {
        cnt++;
        accum(ri, Tt);
      }
#pragma unroll
      for (int e = 0; e < NQ; e++) I[e] += Tt[e];
    }
#pragma unroll
    for (int e = 0; e < NQ; e++) {
      if (done & (1ull << e)) continue;
      double err;
      if (j == 0) {
        p[e] += (I[e] - q[e]);
        err = 16 * fabs(q[e] - 3 * p[e] / 2) / (3 * nn);
      } else {
        err = 16 * fabs(I[e] - 2 * q[e]) / (3 * nn);
        q[e] = I[e];
      }
      if (cnt > 0 && err < T.tol) {
        done |= (1ull << e);
        p[e] = 16 * q[e] / (3 * nn);
      }
    }
    if (done == all) break;
  }
  double *out = qout + U.qoff;
#pragma unroll
  for (int e = 0; e < NQ; e++) out[e] = (done & (1ull << e)) ? p[e] : 0.0;
  const unsigned long long failed = all & ~done;
  qfail[u] = failed;
  if (failed) {
    const int slot = atomicAdd(fbcount, 1);
    fbq[slot] = u;
  }
}

// FM06 linear map of the large grid for zeta_p, P (gc_integrators.c:316-331)
struct Fm06 {
  double i1, i2;
};
__device__ __forceinline__ Fm06 fm06(double zeta_p, double P) {
  const double sigma = 1.0 / sqrt(zeta_p);
  const double t = P - 7.0 * sigma;
  const double rmin = (t > 0.0) ? t : 0.0;
  const double rmax = P + 9.0 * sigma;
  Fm06 m;
  m.i1 = 0.5 * (rmax - rmin);
  m.i2 = 0.5 * (rmax + rmin);
  return m;
}

// type-1 fallback: PSM92 on the FM06 grid of one primitive pair for its failed
// Q entries (type1.c:205-252)
template <int LAB>
__global__ __launch_bounds__(64) void k_fb1(DevTables T, DevJob J, const T1Unit *__restrict__ units,
                                            double *__restrict__ qout, const unsigned long long *__restrict__ qfail,
                                            const int *__restrict__ fbq, const int *__restrict__ fbcount,
                                            unsigned char *__restrict__ status) {
  constexpr int NQ = nq_of(LAB);
  const int count = *fbcount;
  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < count; it += gridDim.x * blockDim.x) {
    const int u = fbq[it];
    const T1Unit U = units[u];
    const unsigned long long mask = qfail[u];
    const ecp_triplet tr = J.trip[U.trip];
    const ecp_centre Cn = J.cent[tr.centre];
    T1Geom G;
    t1_geom(J, tr, U.pa, U.pb, G);
    const double Cc = G.ca * G.cb;
    const double zeta_p = G.za + G.zb;
    const Fm06 M = fm06(zeta_p, (G.za * G.dAC + G.zb * G.dBC) / zeta_p);
    const double minExp = T.lnAcc1;
    const int order = T.lorder;
    // screening break (type1.c:215-225): first r > max(dAC, dBC) with the
    // exponent below the numerical zero
    int brk = order;
    for (int n = 0; n < order; n++) {
      const double r = M.i1 * T.xl[n] + M.i2;
      const double ex = (G.z * r + G.s) * r + G.zd2;
      if (r > G.dAC && r > G.dBC && ex < minExp) {
        brk = n;
        break;
      }
    }
    const int end = brk - 1;
    const ecp_gauss *gs = J.gauss + Cn.g0;
    auto accum = [&](int n, double (&acc)[NQ]) {
      if (n >= brk) return;
      const double r = M.i1 * T.xl[n] + M.i2;
      const double expo = (G.z * r + G.s) * r + G.zd2;
      if (!(expo >= minExp)) return;
      const double g = exp(expo);
      double K[LAB + 1];
      bessel_k<LAB>(T.bessel, G.s * r, K);
      const double Un = eval_ecp(gs, Cn.ng, Cn.L, r);
      const double wn = T.wl[n] * M.i1;
      double rn = 1.0;
      int e = 0;
#pragma unroll
      for (int l1 = 0; l1 <= LAB; l1++) {
        const double a = (Cc * rn) * Un;
#pragma unroll
        for (int l2 = l1; l2 >= 0; l2 -= 2, e++) acc[e] += wn * ((a * K[l2]) * g);
        rn *= r;
      }
    };
    double I[NQ], p[NQ], Tt[NQ];
#pragma unroll
    for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
    accum(T.s92.centre, Tt);
#pragma unroll
    for (int e = 0; e < NQ; e++) {
      I[e] = Tt[e];
      p[e] = I[e];
    }
    unsigned long long done = ~mask;
    for (int lv = 1; lv < T.s92.nlevels && done != ~0ull; lv++) {
      double q[NQ];
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        q[e] = 2 * p[e];
        p[e] = 2 * I[e];
      }
      int cnt = 0;
      for (int k = T.s92.lvl_begin[lv]; k < T.s92.lvl_begin[lv + 1]; k++) {
        const int li = T.p92l[k], ri = T.p92r[k];
#pragma unroll
        for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
        cnt++;
        accum(li, Tt);
        if (ri <= end) {
          cnt++;
          accum(ri, Tt);
        }
#pragma unroll
        for (int e = 0; e < NQ; e++) I[e] += Tt[e];
      }
      const double N = T.s92.lvl_n[lv];
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        if (done & (1ull << e)) continue;
        const double ee = I[e] - p[e];
        if (cnt > 0 && 16 * ee * ee <= 3 * N * fabs(I[e] - q[e]) * T.tol) {
          done |= (1ull << e);
          qout[U.qoff + e] = 16 * I[e] / (3 * N);
        }
      }
    }
    if (done != ~0ull) status[2 * U.trip] = 1;
  }
}

// ---------------------------------------------------------------------------
// type 2: one lane per T entry

struct T2Group {
  int trip, l, eoff, ne, tmpl, lmb;
};

__device__ __forceinline__ int find_group(const int *__restrict__ goff, int ng, int gid) {
  int lo = 0, hi = ng;  // goff[lo] <= gid < goff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (goff[mid] <= gid) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(128) void k_type2(DevTables T, DevJob J, const T2Group *__restrict__ groups,
                                               const int *__restrict__ goff, int ngroups, int ntotal,
                                               const int *__restrict__ tmpl, double *__restrict__ Tout,
                                               unsigned char *__restrict__ tfail, int *__restrict__ gflag,
                                               int *__restrict__ fbq, int *__restrict__ fbcount, int qstride) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ntotal) return;
  const int g = find_group(goff, ngroups, gid);
  const T2Group G = groups[g];
  const int code = tmpl[G.tmpl + (gid - G.eoff)];
  const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
  const ecp_triplet tr = J.trip[G.trip];
  const ecp_centre C = J.cent[tr.centre];
  const double *__restrict__ Fa = J.F + C.F_off + ((size_t)(tr.sa - C.shell0) * J.FL + l1) * SO;
  const double *__restrict__ Fb = J.F + C.F_off + ((size_t)(tr.sb - C.shell0) * J.FL + l2) * SO;
  const double *__restrict__ Uu = J.Ut + C.Ut_off + ((size_t)G.l * J.UK + l3) * SO;
  const int gs = tr.gs, ge = tr.ge;
  auto f = [&](int n) { return T.ws[n] * ((Fa[n] * Fb[n]) * Uu[n]); };
  double p = f(T.s93.centre);
  double q = f(T.s93.first_l) + f(T.s93.first_r);
  double I = p + q;
  bool ok = false;
  double res = 0.0;
  for (int lv = 2; lv < T.s93.nlevels; lv++) {
    const int j = T.s93.lvl_j[lv], nn = T.s93.lvl_n[lv];
    int cnt = 0;
    for (int k = T.s93.lvl_begin[lv]; k < T.s93.lvl_begin[lv + 1]; k++) {
      const int li = T.p93l[k], ri = T.p93r[k];
      double t = 0.0;
      if (li >= gs) {
        cnt++;
        t += f(li);
      }
      if (ri <= ge) {
        cnt++;
        t += f(ri);
      }
      I += t;
    }
    double err;
    if (j == 0) {
      p += (I - q);
      err = 16 * fabs(q - 3 * p / 2) / (3 * nn);
    } else {
      err = 16 * fabs(I - 2 * q) / (3 * nn);
      q = I;
    }
    if (cnt > 0 && err < T.tol) {
      ok = true;
      res = 16 * q / (3 * nn);
      break;
    }
  }
  Tout[gid] = res;
  tfail[gid] = ok ? 0 : 1;
  if (!ok && atomicOr(&gflag[g], 1) == 0) {
    const int slot = atomicAdd(&fbcount[G.lmb], 1);
    fbq[(size_t)G.lmb * qstride + slot] = g;
  }
}

// type-2 fallback: PSM92 on the FM06 grid of one primitive pair for up to
// kFbChunk failed entries of a (triplet, l) group (calcT_FM06, type2.c:417-528)
template <int LMB>
__global__ __launch_bounds__(64) void k_fb2(DevTables T, DevJob J, const T2Group *__restrict__ groups,
                                            const int *__restrict__ tmpl, const unsigned char *__restrict__ tfail,
                                            const int *__restrict__ fbq, const int *__restrict__ fbcount,
                                            int maxChunks, double *__restrict__ fbT,
                                            unsigned char *__restrict__ status) {
  const int count = *fbcount;
  const int per = J.maxPairs * maxChunks;
  const long total = (long)count * per;
  for (long it = blockIdx.x * (long)blockDim.x + threadIdx.x; it < total; it += (long)gridDim.x * blockDim.x) {
    const int g = fbq[it / per];
    const int rem = (int)(it % per);
    const int pair = rem / maxChunks, chunk = rem % maxChunks;
    const T2Group G = groups[g];
    const ecp_triplet tr = J.trip[G.trip];
    const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
    if (pair >= Sa.K * Sb.K) continue;
    // the chunk's failed entries
    int ent[kFbChunk];
    int ne = 0, seen = 0;
    for (int e = 0; e < G.ne && ne < kFbChunk; e++) {
      if (!tfail[G.eoff + e]) continue;
      if (seen++ >= chunk * kFbChunk) ent[ne++] = e;
    }
    if (ne == 0) continue;
    const int pa = pair / Sb.K, pb = pair % Sb.K;
    const ecp_centre Cn = J.cent[tr.centre];
    const ecp_xyz rc = J.atoms[Cn.atom];
    const double dAC = dist(rc, J.atoms[Sa.atom]), dBC = dist(rc, J.atoms[Sb.atom]);
    const double za = J.pa[Sa.prim + pa], zb = J.pa[Sb.prim + pb];
    const double s1 = 2.0 * za * dAC, s2 = 2.0 * zb * dBC;
    const double Cc = J.pd[Sa.prim + pa] * J.pd[Sb.prim + pb];
    const double zeta_p = za + zb;
    const Fm06 M = fm06(zeta_p, (za * dAC + zb * dBC) / zeta_p);
    const double minExp = T.lnAcc2;
    const int order = T.lorder;
    int brk = order;
    for (int n = 0; n < order; n++) {
      const double r = M.i1 * T.xl[n] + M.i2;
      const double d1 = dAC - r, d2 = dBC - r;
      const double ex = -za * d1 * d1 - zb * d2 * d2;
      if (r > dAC && r > dBC && ex < minExp) {
        brk = n;
        break;
      }
    }
    const int end = brk - 1;
    int c1[kFbChunk], c2[kFbChunk], c3[kFbChunk];
    for (int i = 0; i < kFbChunk; i++) {
      const int code = (i < ne) ? tmpl[G.tmpl + ent[i]] : 0;
      c1[i] = code & 0xff;
      c2[i] = (code >> 8) & 0xff;
      c3[i] = (code >> 16) & 0xff;
    }
    const ecp_gauss *gg = J.gauss + Cn.g0;
    auto accum = [&](int n, double (&acc)[kFbChunk]) {
      if (n >= brk) return;
      const double r = M.i1 * T.xl[n] + M.i2;
      const double d1 = dAC - r, d2 = dBC - r;
      const double expo = -za * d1 * d1 - zb * d2 * d2;
      if (!(expo >= minExp)) return;
      const double Un = eval_ecp(gg, Cn.ng, G.l, r);
      double Ka[LMB + 1], Kb[LMB + 1];
      bessel_k<LMB>(T.bessel, s1 * r, Ka);
      bessel_k<LMB>(T.bessel, s2 * r, Kb);
      const double ex = exp(expo);
      const double wn = T.wl[n] * M.i1;
      const double cu = Cc * Un;
#pragma unroll
      for (int i = 0; i < kFbChunk; i++) {
        if (i >= ne) break;
        double rn = 1.0;
        for (int k = 0; k < c3[i]; k++) rn = r * rn;
        double ka = 0.0, kb = 0.0;
#pragma unroll
        for (int l = 0; l <= LMB; l++) {
          if (l == c1[i]) ka = Ka[l];
          if (l == c2[i]) kb = Kb[l];
        }
        acc[i] += wn * ((((cu * rn) * ka) * kb) * ex);
      }
    };
    double I[kFbChunk], p[kFbChunk], Tt[kFbChunk];
    for (int i = 0; i < kFbChunk; i++) Tt[i] = 0.0;
    accum(T.s92.centre, Tt);
    for (int i = 0; i < kFbChunk; i++) {
      I[i] = Tt[i];
      p[i] = I[i];
    }
    unsigned int done = (ne >= 32) ? 0u : ~((1u << ne) - 1u);
    for (int lv = 1; lv < T.s92.nlevels && done != ~0u; lv++) {
      double q[kFbChunk];
      for (int i = 0; i < kFbChunk; i++) {
        q[i] = 2 * p[i];
        p[i] = 2 * I[i];
      }
      int cnt = 0;
      for (int k = T.s92.lvl_begin[lv]; k < T.s92.lvl_begin[lv + 1]; k++) {
        const int li = T.p92l[k], ri = T.p92r[k];
        for (int i = 0; i < kFbChunk; i++) Tt[i] = 0.0;
        cnt++;
        accum(li, Tt);
        if (ri <= end) {
          cnt++;
          accum(ri, Tt);
        }
        for (int i = 0; i < kFbChunk; i++) I[i] += Tt[i];
      }
      const double N = T.s92.lvl_n[lv];
      for (int i = 0; i < kFbChunk; i++) {
        if (done & (1u << i)) continue;
        const double ee = I[i] - p[i];
        if (cnt > 0 && 16 * ee * ee <= 3 * N * fabs(I[i] - q[i]) * T.tol) {
          done |= (1u << i);
          fbT[(size_t)(G.eoff + ent[i]) * J.maxPairs + pair] = 16 * I[i] / (3 * N);
        }
      }
    }
    if (done != ~0u) status[2 * G.trip + 1] = 1;
  }
}

// ---------------------------------------------------------------------------
// assembly: type-1 block (calcChi), type-2 block (calcGamma), binomial shift
// (calcPolynomials) and norms (libecp.c:234-235)

constexpr int kAsmThreads = 64;
constexpr int kMaxC = ECP_DIM_C(ECP_MAX_LBS);           // 56
constexpr int kMaxCL = ECP_DIM_C(2 * ECP_MAX_LBS);      // monomials of degree <= 10
constexpr int kMaxT2 = (ECP_MAX_LBS + ECP_MAX_L) * (ECP_MAX_LBS + ECP_MAX_L) * (2 * ECP_MAX_LBS + 1);

struct AsmArgs {
  const T1Unit *units1;
  const int *t1first;        // per triplet: first type-1 unit
  const double *qout;
  const T2Group *groups;
  const int *tgfirst;        // per triplet: first group
  const int *tmpl;
  const double *Tout;
  const unsigned char *tfail;
  const double *fbT;
  const double *fac, *dfac;
  double *blocks;
};

// binomial shift of a C_DIM(la) x C_DIM(lb) block (in LDS) to the
// IJK(la) x IJK(lb) output (util.c:246-334)
__device__ void shift_block(const DevTables &T, const double *gam, int la, int lb, const double rA[3],
                            const double rB[3], double norm, double *Jbuf, double *out) {
  const int D = T.ijkDim;
  const int ncb = ECP_DIM_C(lb), nia = ECP_DIM_IJK(la), nib = ECP_DIM_IJK(lb);
  const double tol = T.accuracy;
  for (int idx = threadIdx.x; idx < nia * ncb; idx += blockDim.x) {
    const int c1 = idx / ncb, q = idx % ncb;
    const int *e = T.ijk + 3 * ECP_IDX_C(la, c1);
    const int ax = e[0], ay = e[1], az = e[2];
    double acc = 0.0;
    for (int alx = 0; alx <= ax; alx++) {
      const double bx = T.binom[ax * 16 + alx];
      for (int aly = 0; aly <= ay; aly++) {
        const double by = bx * T.binom[ay * 16 + aly];
        for (int alz = 0; alz <= az; alz++) {
          const double bz = by * T.binom[az * 16 + alz];
          // (C - A)^(a - alpha) as unitarySpherePolynomials (util.c:214-243)
          double x = 1.0, y = 1.0, z = 1.0;
          for (int i = 0; i < ax - alx; i++) x = -x * rA[0];
          for (int i = 0; i < ay - aly; i++) y = -y * rA[1];
          for (int i = 0; i < az - alz; i++) z = -z * rA[2];
          const double factor = bz * ((x * y) * z);
          if (fabs(factor) <= tol) continue;
          const int r = T.ijkIndex[(alx * D + aly) * D + alz];
          acc += factor * gam[r * ncb + q];
        }
      }
    }
    Jbuf[idx] = acc;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < nia * nib; idx += blockDim.x) {
    const int c1 = idx / nib, c2 = idx % nib;
    const int *e = T.ijk + 3 * ECP_IDX_C(lb, c2);
    const int bx_ = e[0], by_ = e[1], bz_ = e[2];
    double acc = 0.0;
    for (int btx = 0; btx <= bx_; btx++) {
      const double fx = T.binom[bx_ * 16 + btx];
      for (int bty = 0; bty <= by_; bty++) {
        const double fy = fx * T.binom[by_ * 16 + bty];
        for (int btz = 0; btz <= bz_; btz++) {
          const double fz = fy * T.binom[bz_ * 16 + btz];
          double x = 1.0, y = 1.0, z = 1.0;
          for (int i = 0; i < bx_ - btx; i++) x = -x * rB[0];
          for (int i = 0; i < by_ - bty; i++) y = -y * rB[1];
          for (int i = 0; i < bz_ - btz; i++) z = -z * rB[2];
          double factor = fz * ((x * y) * z);
          if (fabs(factor) <= tol) continue;
          factor *= norm;
          const int r = T.ijkIndex[(btx * D + bty) * D + btz];
          acc += factor * Jbuf[c1 * ncb + r];
        }
      }
    }
    out[idx] = acc;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kAsmThreads) void k_assemble(DevTables T, DevJob J, AsmArgs A, unsigned char *status) {
  __shared__ double blk[kMaxC * kMaxC];
  __shared__ double Jbuf[ECP_DIM_IJK(ECP_MAX_LBS) * kMaxC];
  __shared__ double rsph[(2 * ECP_MAX_LBS + 1) * (2 * ECP_MAX_LBS + 1)];
  __shared__ double Pw[(2 * ECP_MAX_LBS + 1) * (2 * ECP_MAX_LBS + 1)];
  __shared__ double sw[2 * ECP_MAX_LBS + 3], cw[2 * ECP_MAX_LBS + 3];
  __shared__ double fct[kMaxCL * (ECP_MAX_LBS + 1)];
  __shared__ double Tl[kMaxT2];
  const int t = blockIdx.x;
  const ecp_triplet tr = J.trip[t];
  const ecp_centre Cn = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const int la = Sa.l, lb = Sb.l, lab = la + lb;
  const int nca = ECP_DIM_C(la), ncb = ECP_DIM_C(lb);
  const ecp_xyz rc = J.atoms[Cn.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const double rAC[3] = {ra.x - rc.x, ra.y - rc.y, ra.z - rc.z};
  const double rBC[3] = {rb.x - rc.x, rb.y - rc.y, rb.z - rc.z};
  const int D = T.ijkDim;
  const int tmd = ECP_DIM_L(T.tmDim);
  const bool fail1 = status[2 * t] != 0, fail2 = status[2 * t + 1] != 0;

  // ---- type 1 ----
  if (!fail1) {
    for (int i = threadIdx.x; i < nca * ncb; i += blockDim.x) blk[i] = 0.0;
    const int NQ = nq_of(lab);
    int qstart[2 * ECP_MAX_LBS + 1];
    for (int l1 = 0, s = 0; l1 <= lab; l1++) {
      qstart[l1] = s;
      s += l1 / 2 + 1;
    }
    const int u0 = A.t1first[t];
    for (int pr = 0; pr < Sa.K * Sb.K; pr++) {
      const T1Unit U = A.units1[u0 + pr];
      if (threadIdx.x == 0) {
        const double za = J.pa[Sa.prim + U.pa], zb = J.pa[Sb.prim + U.pb];
        double P[3];
        for (int p = 0; p < 3; p++) P[p] = 2.0 * (za * rAC[p] + zb * rBC[p]);
        double r, th, ph;
        sph_coords(P[0], P[1], P[2], r, th, ph);
        real_sph(lab, th, ph, A.fac, A.dfac, rsph, Pw, sw, cw);
      }
      __syncthreads();
      // factor[p][k] = sum_m S_lm(k^) poly2sph[p][lm], l = deg(p) - 2k
      const int np = ECP_DIM_C(lab);
      const int kd = lab / 2 + 1;
      for (int idx = threadIdx.x; idx < np * kd; idx += blockDim.x) {
        const int p = idx / kd, k = idx % kd;
        const int *e = T.ijk + 3 * p;
        const int deg = e[0] + e[1] + e[2];
        const int l = deg - 2 * k;
        double f = 0.0;
        if (l >= 0)
          for (int m = 0; m < 2 * l + 1; m++) f += rsph[ECP_IDX_LM(l, m)] * T.poly2sph[(size_t)p * tmd + ECP_IDX_LM(l, m)];
        fct[idx] = f;
      }
      __syncthreads();
      const double *Q = A.qout + U.qoff;
      for (int idx = threadIdx.x; idx < nca * ncb; idx += blockDim.x) {
        const int i = idx / ncb, j = idx % ncb;
        const int *ea = T.ijk + 3 * i, *eb = T.ijk + 3 * j;
        const int lx = ea[0] + eb[0], ly = ea[1] + eb[1], lz = ea[2] + eb[2];
        const int lmax = lx + ly + lz;
        const int p = T.ijkIndex[(lx * D + ly) * D + lz];
        double acc = blk[idx];
        for (int l = lmax; l >= 0; l -= 2)
          acc += fct[p * kd + (lmax - l) / 2] * Q[qstart[lmax] + (lmax - l) / 2];
        blk[idx] = acc;
      }
      __syncthreads();
    }
    (void)NQ;
    const double norm1 = 4.0 * M_PI;
    shift_block(T, blk, la, lb, rAC, rBC, norm1, Jbuf, A.blocks + tr.blk1);
  }

  // ---- type 2 ----
  if (!fail2) {
    for (int i = threadIdx.x; i < nca * ncb; i += blockDim.x) blk[i] = 0.0;
    __syncthreads();
    const int nlm2 = J.angL * J.angL;
    const double *angA = J.ang + Cn.ang_off +
                         (size_t)(Sa.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
    const double *angB = J.ang + Cn.ang_off +
                         (size_t)(Sb.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
    const int g0 = A.tgfirst[t];
    for (int l = 0; l < Cn.L; l++) {
      const T2Group G = A.groups[g0 + l];
      const int d1 = la + l + 1, d2 = lb + l + 1, d3 = lab + 1;
      for (int i = threadIdx.x; i < d1 * d2 * d3; i += blockDim.x) Tl[i] = 0.0;
      __syncthreads();
      for (int e = threadIdx.x; e < G.ne; e += blockDim.x) {
        const int code = A.tmpl[G.tmpl + e];
        const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
        double v;
        if (!A.tfail[G.eoff + e]) {
          v = A.Tout[G.eoff + e];
        } else {
          v = 0.0;
          const double *fb = A.fbT + (size_t)(G.eoff + e) * J.maxPairs;
          for (int pr = 0; pr < Sa.K * Sb.K; pr++) v += fb[pr];
        }
        Tl[(l1 * d2 + l2) * d3 + l3] = v;
      }
      __syncthreads();
      for (int idx = threadIdx.x; idx < nca * ncb; idx += blockDim.x) {
        const int pp = idx / ncb, qq = idx % ncb;
        // alpha, beta of the C-centred monomials pp, qq (C_INDEX order)
        int alpha = 0, beta = 0;
        while (ECP_DIM_C(alpha) <= pp) alpha++;
        while (ECP_DIM_C(beta) <= qq) beta++;
        const int par1 = (alpha + l) % 2, par2 = (beta + l) % 2;
        const int ll1 = (par1 > l - alpha) ? par1 : l - alpha;
        const int ll2 = (par2 > l - beta) ? par2 : l - beta;
        double tmp = 0.0;
        for (int lam1 = ll1; lam1 <= la + l; lam1 += 2)
          for (int lam2 = ll2; lam2 <= lb + l; lam2 += 2) {
            double factor = 0.0;
            for (int m = 0; m < 2 * l + 1; m++)
              factor += angA[((size_t)lam1 * nlm2 + ECP_IDX_LM(l, m)) * J.ang_c + pp] *
                        angB[((size_t)lam2 * nlm2 + ECP_IDX_LM(l, m)) * J.ang_c + qq];
            tmp += factor * Tl[(lam1 * d2 + lam2) * d3 + alpha + beta];
          }
        blk[idx] += tmp;
      }
      __syncthreads();
    }
    const double norm2 = (4.0 * M_PI) * (4.0 * M_PI);
    shift_block(T, blk, la, lb, rAC, rBC, norm2, Jbuf, A.blocks + tr.blk2);
  }
}

// ---------------------------------------------------------------------------
// host-side orchestration

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct ecp_dev {
  hipStream_t stream = nullptr;
  DevTables T{};
  // static device tables
  void *static_mem = nullptr;
  // job arena (grow-only)
  char *arena = nullptr;
  size_t arena_cap = 0;
  char *pin_in = nullptr;
  size_t pin_in_cap = 0;
  char *pin_out = nullptr;
  size_t pin_out_cap = 0;
  double *dfac_dev = nullptr, *fac_dev = nullptr;
  hipEvent_t ev[8];
  float last_ms[8];
  int nev = 0;
  // template tables (host)
  std::vector<int> tmpl;
  std::vector<int> tmpl_off;  // index (la*16 + lb)*16 + l -> offset, count in tmpl_cnt
  std::vector<int> tmpl_cnt;
  int *tmpl_dev = nullptr;
};

namespace {

int tmpl_key(int la, int lb, int l) { return (la * 16 + lb) * 16 + l; }

// T_l entries (lambda1, lambda2, alpha+beta) with a possibly non-zero angular
// factor: lambda1 from max((alpha+l) mod 2, l-alpha) to l+alpha in steps of 2
// (the lower limit of type2.c:590-592; the angular table vanishes above
// l+alpha, angular_integrals.c:77), likewise lambda2 with beta.
void build_templates(ecp_dev *d) {
  d->tmpl.clear();
  d->tmpl_off.assign(16 * 16 * 16, 0);
  d->tmpl_cnt.assign(16 * 16 * 16, 0);
  for (int la = 0; la <= ECP_MAX_LBS; la++)
    for (int lb = 0; lb <= ECP_MAX_LBS; lb++)
      for (int l = 0; l < ECP_MAX_L; l++) {
        const int key = tmpl_key(la, lb, l);
        d->tmpl_off[key] = (int)d->tmpl.size();
        const int d1 = la + l + 1, d2 = lb + l + 1, d3 = la + lb + 1;
        std::vector<char> used((size_t)d1 * d2 * d3, 0);
        for (int alpha = 0; alpha <= la; alpha++)
          for (int beta = 0; beta <= lb; beta++)
            for (int l1 = std::max((alpha + l) % 2, l - alpha); l1 <= l + alpha; l1 += 2)
              for (int l2 = std::max((beta + l) % 2, l - beta); l2 <= l + beta; l2 += 2)
                used[((size_t)l1 * d2 + l2) * d3 + alpha + beta] = 1;
        for (int l1 = 0; l1 < d1; l1++)
          for (int l2 = 0; l2 < d2; l2++)
            for (int l3 = 0; l3 < d3; l3++)
              if (used[((size_t)l1 * d2 + l2) * d3 + l3]) d->tmpl.push_back(l1 | (l2 << 8) | (l3 << 16));
        d->tmpl_cnt[key] = (int)d->tmpl.size() - d->tmpl_off[key];
      }
}

}  // namespace

extern "C" ecp_dev *ecp_dev_create(const ecp_static *st, char *err, size_t errlen) {
  ecp_dev *d = new ecp_dev();
  auto fail = [&](const char *what, hipError_t e) {
    snprintf(err, errlen, "%s: %s", what, hipGetErrorString(e));
    ecp_dev_destroy(d);
    return (ecp_dev *)nullptr;
  };
  hipError_t e;
  int ndev = 0;
  if ((e = hipGetDeviceCount(&ndev)) != hipSuccess || ndev == 0) {
    snprintf(err, errlen, "no HIP device available (%s)", hipGetErrorString(e));
    delete d;
    return nullptr;
  }
  if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) return fail("hipStreamCreate", e);
  for (int i = 0; i < 8; i++)
    if ((e = hipEventCreate(&d->ev[i])) != hipSuccess) return fail("hipEventCreate", e);
  build_templates(d);
  // static tables in one allocation
  const int so = ECP_SMALL_ORDER, lo = st->lorder;
  const size_t nbess = (size_t)(ECP_BESSEL_N + 1) * ECP_BESSEL_STRIDE;
  const size_t npoly = (size_t)ECP_DIM_C(st->tmDim) * ECP_DIM_L(st->tmDim);
  const size_t nom = (size_t)ECP_DIM_L(st->omL) * ECP_DIM_L(st->omLambda) * ECP_DIM_C(st->omAlpha);
  const size_t nijk = 3 * (size_t)ECP_DIM_C(st->ijkDim - 1);
  const size_t nidx = (size_t)st->ijkDim * st->ijkDim * st->ijkDim;
  const size_t nfac = 2 * (size_t)st->tmDim + 6;
  std::vector<double> binom(16 * 16, 0.0);
  for (int n = 0; n < 16; n++)
    for (int k = 0; k <= n; k++) {
      double f[16];
      f[0] = 1.0;
      for (int i = 1; i < 16; i++) f[i] = f[i - 1] * i;
      binom[n * 16 + k] = f[n] / (f[n - k] * f[k]);
    }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align_up(bytes);
    return o;
  };
  const size_t o_xs = take(so * 8), o_ws = take(so * 8), o_xl = take(lo * 8), o_wl = take(lo * 8);
  const size_t o_p93l = take(so * 4), o_p93r = take(so * 4), o_p92l = take(lo * 4), o_p92r = take(lo * 4);
  const size_t o_bess = take(nbess * 8), o_poly = take(npoly * 8), o_om = take(nom * 8);
  const size_t o_ijk = take(nijk * 4), o_idx = take(nidx * 4), o_bin = take(256 * 8);
  const size_t o_fac = take(nfac * 8), o_dfac = take(nfac * 8);
  const size_t o_tmpl = take(d->tmpl.size() * 4 + 4);
  std::vector<char> h(off, 0);
  auto put = [&](size_t o, const void *src, size_t bytes) { memcpy(h.data() + o, src, bytes); };
  put(o_xs, st->xs, so * 8);
  put(o_ws, st->ws, so * 8);
  put(o_xl, st->xl, lo * 8);
  put(o_wl, st->wl, lo * 8);
  put(o_p93l, st->p93l, so * 4);
  put(o_p93r, st->p93r, so * 4);
  put(o_p92l, st->p92l, lo * 4);
  put(o_p92r, st->p92r, lo * 4);
  put(o_bess, st->bessel, nbess * 8);
  put(o_poly, st->poly2sph, npoly * 8);
  put(o_om, st->omega, nom * 8);
  put(o_ijk, st->ijk, nijk * 4);
  put(o_idx, st->ijkIndex, nidx * 4);
  put(o_bin, binom.data(), 256 * 8);
  put(o_fac, st->fac, nfac * 8);
  put(o_dfac, st->dfac, nfac * 8);
  put(o_tmpl, d->tmpl.data(), d->tmpl.size() * 4);
  if ((e = hipMalloc(&d->static_mem, off)) != hipSuccess) return fail("hipMalloc(static tables)", e);
  if ((e = hipMemcpy(d->static_mem, h.data(), off, hipMemcpyHostToDevice)) != hipSuccess)
    return fail("hipMemcpy(static tables)", e);
  char *b = (char *)d->static_mem;
  DevTables &T = d->T;
  T.xs = (const double *)(b + o_xs);
  T.ws = (const double *)(b + o_ws);
  T.xl = (const double *)(b + o_xl);
  T.wl = (const double *)(b + o_wl);
  T.p93l = (const int *)(b + o_p93l);
  T.p93r = (const int *)(b + o_p93r);
  T.p92l = (const int *)(b + o_p92l);
  T.p92r = (const int *)(b + o_p92r);
  T.bessel = (const double *)(b + o_bess);
  T.poly2sph = (const double *)(b + o_poly);
  T.omega = (const double *)(b + o_om);
  T.ijk = (const int *)(b + o_ijk);
  T.ijkIndex = (const int *)(b + o_idx);
  T.binom = (const double *)(b + o_bin);
  d->fac_dev = (double *)(b + o_fac);
  d->dfac_dev = (double *)(b + o_dfac);
  d->tmpl_dev = (int *)(b + o_tmpl);
  T.ijkDim = st->ijkDim;
  T.tmDim = st->tmDim;
  T.lorder = st->lorder;
  T.omL = st->omL;
  T.omLambda = st->omLambda;
  T.omAlpha = st->omAlpha;
  T.tol = st->tolerance;
  T.lnAcc1 = st->lnAcc1;
  T.lnAcc2 = st->lnAcc2;
  T.accuracy = st->accuracy;
  T.s93 = st->s93;
  T.s92 = st->s92;
  return d;
}

extern "C" void ecp_dev_destroy(ecp_dev *d) {
  if (!d) return;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  if (d->static_mem) (void)hipFree(d->static_mem);
  if (d->arena) (void)hipFree(d->arena);
  if (d->pin_in) (void)hipHostFree(d->pin_in);
  if (d->pin_out) (void)hipHostFree(d->pin_out);
  for (int i = 0; i < 8; i++)
    if (d->ev[i]) (void)hipEventDestroy(d->ev[i]);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

extern "C" int ecp_dev_last_timing(ecp_dev *d, double *ms, int n) {
  int k = std::min(n, d->nev);
  for (int i = 0; i < k; i++) ms[i] = d->last_ms[i];
  return k;
}

namespace {

template <int LAB>
void launch_type1(hipStream_t s, const DevTables &T, const DevJob &J, const T1Unit *u, int n, double *q,
                  unsigned long long *qf, int *fbq, int *fbc) {
  if (n > 0) k_type1<LAB><<<(n + 63) / 64, 64, 0, s>>>(T, J, u, n, q, qf, fbq, fbc);
}

template <int LAB>
void launch_fb1(hipStream_t s, const DevTables &T, const DevJob &J, const T1Unit *u, int n, double *q,
                const unsigned long long *qf, const int *fbq, const int *fbc, unsigned char *st) {
  if (n > 0) {
    const int blocks = std::min((n + 63) / 64, 1024);
    k_fb1<LAB><<<blocks, 64, 0, s>>>(T, J, u, q, qf, fbq, fbc, st);
  }
}

template <int LMB>
void launch_fb2(hipStream_t s, const DevTables &T, const DevJob &J, const T2Group *g, const int *tm,
                const unsigned char *tf, const int *fbq, const int *fbc, int maxChunks, double *fbT,
                unsigned char *st, long maxItems) {
  if (maxItems > 0) {
    const long blocks = std::min<long>((maxItems + 63) / 64, 2048);
    k_fb2<LMB><<<(int)blocks, 64, 0, s>>>(T, J, g, tm, tf, fbq, fbc, maxChunks, fbT, st);
  }
}

}  // namespace

extern "C" int ecp_dev_run(ecp_dev *d, const ecp_job *job, double *blocks_out, unsigned char *status_out,
                           char *err, size_t errlen) {
  const int nt = job->ntrip;
  d->nev = 0;
  if (nt == 0) return 0;
  // ---- derived work lists (host) ----
  std::vector<T1Unit> units1;
  std::vector<int> t1first(nt), cls_begin(ECP_MAX_LAB + 2, 0), cls_count(ECP_MAX_LAB + 1, 0);
  // type-1 units grouped by class lab, inside a class in triplet order
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    cls_count[job->shells[tr.sa].l + job->shells[tr.sb].l] += job->shells[tr.sa].K * job->shells[tr.sb].K;
  }
  for (int c = 0; c <= ECP_MAX_LAB; c++) cls_begin[c + 1] = cls_begin[c] + cls_count[c];
  units1.resize(cls_begin[ECP_MAX_LAB + 1]);
  std::vector<int> cls_fill(cls_begin.begin(), cls_begin.end() - 1);
  std::vector<size_t> cls_qbase(ECP_MAX_LAB + 2, 0);
  for (int c = 0; c <= ECP_MAX_LAB; c++) cls_qbase[c + 1] = cls_qbase[c] + (size_t)cls_count[c] * nq_of(c);
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    const ecp_shell &Sa = job->shells[tr.sa], &Sb = job->shells[tr.sb];
    const int c = Sa.l + Sb.l;
    t1first[t] = cls_fill[c];
    for (int pa = 0; pa < Sa.K; pa++)
      for (int pb = 0; pb < Sb.K; pb++) {
        const int u = cls_fill[c]++;
        units1[u] = T1Unit{t, pa, pb, (int)(cls_qbase[c] + (size_t)(u - cls_begin[c]) * nq_of(c))};
      }
  }
  const size_t nq = cls_qbase[ECP_MAX_LAB + 1];
  // type-2 groups
  std::vector<T2Group> groups;
  std::vector<int> goff, tgfirst(nt);
  int ntot2 = 0, maxNE = 0;
  int lmb_count[kMaxLMB + 1] = {0};
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    const int la = job->shells[tr.sa].l, lb = job->shells[tr.sb].l;
    tgfirst[t] = (int)groups.size();
    for (int l = 0; l < job->cent[tr.centre].L; l++) {
      const int key = tmpl_key(la, lb, l);
      T2Group g{t, l, ntot2, d->tmpl_cnt[key], d->tmpl_off[key], std::max(la, lb) + l};
      groups.push_back(g);
      goff.push_back(ntot2);
      ntot2 += g.ne;
      maxNE = std::max(maxNE, g.ne);
      lmb_count[g.lmb]++;
    }
  }
  const int ngroups = (int)groups.size();
  goff.push_back(ntot2);
  // F and angular work items
  std::vector<int2> fitems, aitems;
  for (int c = 0; c < job->ncent; c++) {
    const ecp_centre &C = job->cent[c];
    for (int s = 0; s < C.nsh; s++)
      if (!job->win[C.win_off + s].skip) fitems.push_back(make_int2(c, s));
    for (int a = 0; a < C.nat; a++) aitems.push_back(make_int2(c, a));
  }
  const int maxPairs = job->maxK * job->maxK;
  const int maxChunks = (maxNE + kFbChunk - 1) / kFbChunk;
  // ---- input staging (one pinned buffer, one copy) ----
  size_t in_off = 0;
  auto in_take = [&](size_t bytes) {
    size_t o = in_off;
    in_off += align_up(bytes);
    return o;
  };
  const size_t i_atoms = in_take(sizeof(ecp_xyz) * job->nat), i_amaxl = in_take(4 * job->nat);
  const size_t i_sh = in_take(sizeof(ecp_shell) * job->nsh), i_pa = in_take(8 * job->nprim),
               i_pd = in_take(8 * job->nprim);
  const size_t i_cent = in_take(sizeof(ecp_centre) * job->ncent), i_gauss = in_take(sizeof(ecp_gauss) * job->ngauss);
  const size_t i_win = in_take(sizeof(ecp_window) * job->nwin), i_trip = in_take(sizeof(ecp_triplet) * nt);
  const size_t i_u1 = in_take(sizeof(T1Unit) * units1.size()), i_t1f = in_take(4 * (size_t)nt);
  const size_t i_grp = in_take(sizeof(T2Group) * groups.size()), i_goff = in_take(4 * goff.size());
  const size_t i_tgf = in_take(4 * (size_t)nt), i_fit = in_take(sizeof(int2) * fitems.size()),
               i_ait = in_take(sizeof(int2) * aitems.size());
  const size_t in_bytes = in_off;
  // device-only buffers
  size_t dv_off = align_up(in_bytes);
  auto dv_take = [&](size_t bytes) {
    size_t o = dv_off;
    dv_off += align_up(bytes);
    return o;
  };
  const size_t d_UL = dv_take(8 * job->nUL), d_Ut = dv_take(8 * job->nUt), d_F = dv_take(8 * job->nF),
               d_ang = dv_take(8 * job->nang);
  const size_t d_q = dv_take(8 * nq), d_qf = dv_take(8 * units1.size());
  const size_t d_fbq1 = dv_take(4 * units1.size()), d_fbc = dv_take(4 * 64);
  const size_t d_T = dv_take(8 * (size_t)ntot2), d_tf = dv_take((size_t)ntot2);
  const size_t d_gflag = dv_take(4 * (size_t)ngroups);
  const size_t d_fbq2 = dv_take(4 * (size_t)(kMaxLMB + 1) * ngroups);
  const size_t d_fbT = dv_take(8 * (size_t)ntot2 * maxPairs);
  const size_t d_blk = dv_take(8 * job->nblk), d_stat = dv_take(2 * (size_t)nt);
  const size_t total = dv_off;
  if (total > d->arena_cap) {
    if (d->arena) ECP_CHECK(hipFree(d->arena));
    d->arena = nullptr;
    d->arena_cap = 0;
    const size_t cap = total + total / 4;
    ECP_CHECK(hipMalloc(&d->arena, cap));
    d->arena_cap = cap;
  }
  if (in_bytes > d->pin_in_cap) {
    if (d->pin_in) ECP_CHECK(hipHostFree(d->pin_in));
    d->pin_in = nullptr;
    ECP_CHECK(hipHostMalloc(&d->pin_in, in_bytes + in_bytes / 4));
    d->pin_in_cap = in_bytes + in_bytes / 4;
  }
  const size_t out_bytes = 8 * job->nblk + 2 * (size_t)nt;
  if (out_bytes > d->pin_out_cap) {
    if (d->pin_out) ECP_CHECK(hipHostFree(d->pin_out));
    d->pin_out = nullptr;
    ECP_CHECK(hipHostMalloc(&d->pin_out, out_bytes + out_bytes / 4));
    d->pin_out_cap = out_bytes + out_bytes / 4;
  }
  char *P = d->pin_in;
  auto cp = [&](size_t o, const void *src, size_t bytes) {
    if (bytes) memcpy(P + o, src, bytes);
  };
  cp(i_atoms, job->atoms, sizeof(ecp_xyz) * job->nat);
  cp(i_amaxl, job->atom_maxl, 4 * (size_t)job->nat);
  cp(i_sh, job->shells, sizeof(ecp_shell) * job->nsh);
  cp(i_pa, job->pa, 8 * (size_t)job->nprim);
  cp(i_pd, job->pd, 8 * (size_t)job->nprim);
  cp(i_cent, job->cent, sizeof(ecp_centre) * job->ncent);
  cp(i_gauss, job->gauss, sizeof(ecp_gauss) * job->ngauss);
  cp(i_win, job->win, sizeof(ecp_window) * job->nwin);
  cp(i_trip, job->trip, sizeof(ecp_triplet) * nt);
  cp(i_u1, units1.data(), sizeof(T1Unit) * units1.size());
  cp(i_t1f, t1first.data(), 4 * (size_t)nt);
  cp(i_grp, groups.data(), sizeof(T2Group) * groups.size());
  cp(i_goff, goff.data(), 4 * goff.size());
  cp(i_tgf, tgfirst.data(), 4 * (size_t)nt);
  cp(i_fit, fitems.data(), sizeof(int2) * fitems.size());
  cp(i_ait, aitems.data(), sizeof(int2) * aitems.size());
  char *A = d->arena;
  hipStream_t s = d->stream;
  ECP_CHECK(hipMemcpyAsync(A, P, in_bytes, hipMemcpyHostToDevice, s));
  ECP_CHECK(hipMemsetAsync(A + d_fbc, 0, 4 * 64, s));
  ECP_CHECK(hipMemsetAsync(A + d_gflag, 0, 4 * (size_t)ngroups, s));
  ECP_CHECK(hipMemsetAsync(A + d_stat, 0, 2 * (size_t)nt, s));
  ECP_CHECK(hipMemsetAsync(A + d_fbT, 0, 8 * (size_t)ntot2 * maxPairs, s));
  DevJob J{};
  J.atoms = (const ecp_xyz *)(A + i_atoms);
  J.atom_maxl = (const int *)(A + i_amaxl);
  J.shells = (const ecp_shell *)(A + i_sh);
  J.pa = (const double *)(A + i_pa);
  J.pd = (const double *)(A + i_pd);
  J.cent = (const ecp_centre *)(A + i_cent);
  J.gauss = (const ecp_gauss *)(A + i_gauss);
  J.win = (const ecp_window *)(A + i_win);
  J.trip = (const ecp_triplet *)(A + i_trip);
  J.ntrip = nt;
  J.UL = (double *)(A + d_UL);
  J.Ut = (double *)(A + d_Ut);
  J.F = (double *)(A + d_F);
  J.ang = (double *)(A + d_ang);
  J.FL = job->FL;
  J.UK = job->UK;
  J.angL = job->angL;
  J.ang_lam = job->ang_lam;
  J.ang_c = job->ang_c;
  J.maxPairs = maxPairs;
  const DevTables &T = d->T;
  const T1Unit *u1 = (const T1Unit *)(A + i_u1);
  double *q = (double *)(A + d_q);
  unsigned long long *qf = (unsigned long long *)(A + d_qf);
  int *fbq1 = (int *)(A + d_fbq1);
  int *fbc = (int *)(A + d_fbc);  // [0..15] type-1 queues, [16..31] type-2 queues
  unsigned char *stat = (unsigned char *)(A + d_stat);
  const T2Group *grp = (const T2Group *)(A + i_grp);
  int ei = 0;
  ECP_CHECK(hipEventRecord(d->ev[ei++], s));
  k_utab<<<job->ncent, 384, 0, s>>>(T, J);
  if (!fitems.empty()) k_ftab<<<(int)fitems.size(), 128, 0, s>>>(T, J, (const int2 *)(A + i_fit));
  if (!aitems.empty())
    k_ang<<<(int)aitems.size(), 128, 0, s>>>(T, J, (const int2 *)(A + i_ait), d->fac_dev, d->dfac_dev);
  ECP_CHECK(hipGetLastError());
  ECP_CHECK(hipEventRecord(d->ev[ei++], s));
  // type 1
#define T1(c)                                                                                          \
  launch_type1<c>(s, T, J, u1 + cls_begin[c], cls_count[c], q, qf + cls_begin[c], fbq1 + cls_begin[c], \
                  fbc + c);
  T1(0) T1(1) T1(2) T1(3) T1(4) T1(5) T1(6) T1(7) T1(8) T1(9) T1(10)
#undef T1
  ECP_CHECK(hipGetLastError());
  // type 2
  double *Tout = (double *)(A + d_T);
  unsigned char *tf = (unsigned char *)(A + d_tf);
  int *gflag = (int *)(A + d_gflag);
  int *fbq2 = (int *)(A + d_fbq2);
  if (ntot2 > 0)
    k_type2<<<(ntot2 + 127) / 128, 128, 0, s>>>(T, J, grp, (const int *)(A + i_goff), ngroups, ntot2, d->tmpl_dev,
                                                 Tout, tf, gflag, fbq2, fbc + 16, ngroups);
  ECP_CHECK(hipGetLastError());
  ECP_CHECK(hipEventRecord(d->ev[ei++], s));
  // fallbacks (persistent over the device-built queues)
#define FB1(c)                                                                                        \
  launch_fb1<c>(s, T, J, u1 + cls_begin[c], cls_count[c], q, qf + cls_begin[c], fbq1 + cls_begin[c], \
                fbc + c, stat);
  FB1(0) FB1(1) FB1(2) FB1(3) FB1(4) FB1(5) FB1(6) FB1(7) FB1(8) FB1(9) FB1(10)
#undef FB1
  double *fbT = (double *)(A + d_fbT);
#define FB2(c)                                                                                         \
  launch_fb2<c>(s, T, J, grp, d->tmpl_dev, tf, fbq2 + (size_t)c * ngroups, fbc + 16 + c, maxChunks, fbT, \
                stat, (long)lmb_count[c] * maxPairs * maxChunks);
  FB2(0) FB2(1) FB2(2) FB2(3) FB2(4) FB2(5) FB2(6) FB2(7) FB2(8) FB2(9) FB2(10) FB2(11)
#undef FB2
  ECP_CHECK(hipGetLastError());
  ECP_CHECK(hipEventRecord(d->ev[ei++], s));
  AsmArgs aa;
  aa.units1 = u1;
  aa.t1first = (const int *)(A + i_t1f);
  aa.qout = q;
  aa.groups = grp;
  aa.tgfirst = (const int *)(A + i_tgf);
  aa.tmpl = d->tmpl_dev;
  aa.Tout = Tout;
  aa.tfail = tf;
  aa.fbT = fbT;
  aa.fac = d->fac_dev;
  aa.dfac = d->dfac_dev;
  aa.blocks = (double *)(A + d_blk);
  k_assemble<<<nt, kAsmThreads, 0, s>>>(T, J, aa, stat);
  ECP_CHECK(hipGetLastError());
  ECP_CHECK(hipEventRecord(d->ev[ei++], s));
  ECP_CHECK(hipMemcpyAsync(d->pin_out, A + d_blk, 8 * job->nblk, hipMemcpyDeviceToHost, s));
  ECP_CHECK(hipMemcpyAsync(d->pin_out + 8 * job->nblk, A + d_stat, 2 * (size_t)nt, hipMemcpyDeviceToHost, s));
  ECP_CHECK(hipStreamSynchronize(s));
  for (int i = 1; i < ei; i++) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, d->ev[i - 1], d->ev[i]);
    d->last_ms[i - 1] = ms;
  }
  d->nev = ei - 1;
  memcpy(blocks_out, d->pin_out, 8 * job->nblk);
  memcpy(status_out, d->pin_out + 8 * job->nblk, 2 * (size_t)nt);
  return 0;
}
