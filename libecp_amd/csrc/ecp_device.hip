// HIP device layer of the ECP integral engine, written for gfx950 (MI355X).
//
// Pipeline of one job (all FP64; a main stream forks into worker streams
// and joins them, captured once into a HIP graph):
//   k_tables zero-fill of counters and flags; U_L(r_n) and r_n^k U_l(r_n) on
//            the 383-point small grid per (centre, channel, point chunk)
//            (evalECP / tabECP_FM06: ecp.c:41-60, type2.c:184-219); and
//            F_lambda(r_n) per (centre, shell) on its screened window
//            (calcF_FM06: type2.c:222-310)
//   k_ang    per-(centre, atom) angular tables (evalAngularIntegrals:
//            angular_integrals.c:104-142), concurrent with the quadratures
//   k_type1 / k_t1tail  type-1 radial integrals Q[l1][l2] per (triplet,
//            primitive pair) with the PS93 sequence (calcQ: type1.c:94-203):
//            lane-serial head, wave-cooperative tail for the open pairs
//   k_type2  type-2 radial integrals T_l[l1][l2][l3] per entry with PS93
//            (calcT_fast_FM06: type2.c:336-381)
//   k_fb1 / k_fb2  PSM92 large-grid redo of failed entries on FM06-mapped
//            grids (type1.c:205-252, calcT_FM06: type2.c:417-528)
//   k_asm1 / k_asm2  type-1 / type-2 blocks, binomial shift and norms
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
#include <mutex>
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

// Device scratch of the no-allocation entry point (ecp_dev_contract_noalloc):
// one array per GPU, held until the process ends.  Host code never uses the
// name `scratch` as an address: set-up looks up each device's copy with
// hipGetSymbolAddress, and every use goes through that pointer.
constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices];

// The scratch address on the current device, looked up once per device.
// 0, or -1 (message in err).
int scratch_setup(void **ptr, char *err, size_t errlen) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kScratchDevices) {
    snprintf(err, errlen, "no-allocation scratch: no usable current device (%s)", hipGetErrorString(e));
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      snprintf(err, errlen, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return -1;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

namespace {

constexpr int SO = ECP_SMALL_ORDER;
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
  int deriv;  // derivative order: F slabs per run 0..deriv, angular tables to lmax + deriv
  int ptmajor;  // the type-2 F and U rows point-major (per-class sequence), else row-major
};

// ---------------------------------------------------------------------------
// numerical primitives

// Radial powers other than 0, 1, 2 (rare in ECPs) through a call: inlined,
// pow's registers set the fallbacks' budget (the type-1 fallback 152
// registers and three waves per SIMD, against 128 and four)
__device__ __attribute__((noinline)) double ecp_pow_general(double r, double n) { return pow(r, n); }

__device__ __forceinline__ double ecp_pow(double r, double n) {
  // exact special cases of a correctly rounded pow
  if (n == 2.0) return r * r;
  if (n == 1.0) return r;
  if (n == 0.0) return 1.0;
  return ecp_pow_general(r, n);
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

// Real spherical harmonics written in place into out[(lmax+1)^2] (LDS or
// global), with the arithmetic of spherical_harmonics.c:15-114: associated
// Legendre recurrences without Condon-Shortley phase, normalisation
// sqrt((2l+1)/(2pi) (l-m)!/(l+m)!), 1/sqrt(2) for m = 0, sin/cos(m phi) by
// the angle-addition recurrence.
__device__ void real_sph_inplace(int lmax, double theta, double phi, const double *__restrict__ fac,
                                 const double *__restrict__ dfac, double *out) {
#define RSH(l, m) ((l) * (l) + (l) + (m))
  const double x = cos(theta);
  for (int i = 0; i < (lmax + 1) * (lmax + 1); i++) out[i] = 0.0;
  if (1.0 == x) {
    for (int l = 0; l <= lmax; l++) out[RSH(l, 0)] = 1.0;
  } else if (-1.0 == x) {
    out[0] = 1.0;
    for (int l = 1; l <= lmax; l++) out[RSH(l, 0)] = -out[RSH(l - 1, 0)];
  } else {
    const double s1 = sqrt(1.0 - x * x);
    double sl = 1.0;
    out[0] = 1.0;
    for (int l = 1; l <= lmax; l++) {
      sl = (l == 1) ? s1 : sl * s1;
      out[RSH(l, l)] = sl * dfac[2 * l - 1];
      int m = l - 1;
      out[RSH(l, m)] = x * (2 * m + 1) * out[RSH(l - 1, m)];
      if (l > 1)
        for (m = 0; m <= l - 2; m++)
          out[RSH(l, m)] = (x * (2 * l - 1) * out[RSH(l - 1, m)] - (l + m - 1) * out[RSH(l - 2, m)]) / (l - m);
    }
  }
  const double sp = sin(phi), cp = cos(phi);
  for (int l = 0; l <= lmax; l++) {
    const double norm0 = sqrt((2.0 * l + 1.0) / (2.0 * M_PI));
    out[RSH(l, 0)] = norm0 * out[RSH(l, 0)];
    for (int m = 1; m <= l; m++) out[RSH(l, m)] = (sqrt(fac[l - m] / fac[l + m]) * norm0) * out[RSH(l, m)];
    out[RSH(l, 0)] = out[RSH(l, 0)] / sqrt(2.0);
    double sm = 0.0, cm = 1.0;
    for (int m = 1; m <= l; m++) {
      if (0.0 == phi) {
        sm = 0.0;
        cm = 1.0;
      } else if (m == 1) {
        sm = sp;
        cm = cp;
      } else {
        const double ns = sp * cm + cp * sm;
        cm = cp * cm - sp * sm;
        sm = ns;
      }
      const double P = out[RSH(l, m)];
      out[RSH(l, -m)] = P * sm;
      out[RSH(l, m)] = P * cm;
    }
  }
#undef RSH
}

// Column m >= 0 of real_sph_inplace (entries (l, +-m), l = m..lmax) with
// the same arithmetic, so that the columns of one direction can be formed by
// different lanes: s1^m by repeated multiplication, the diagonal, the first
// subdiagonal and the three-term recurrence in l, the normalisation per
// (l, m), and sin/cos(m phi) by the angle-addition recurrence from m = 1.
// x = cos(theta), sp = sin(phi), cp = cos(phi): the direction's trigonometric
// values, formed once per direction by the caller.
__device__ void real_sph_column(int lmax, int m, double x, double phi, double sp, double cp,
                                const double *__restrict__ fac, const double *__restrict__ dfac, double *out) {
#define RSH(l, mm) ((l) * (l) + (l) + (mm))
  if (1.0 == x || -1.0 == x) {
    for (int l = m; l <= lmax; l++) {
      out[RSH(l, m)] = 0.0;
      out[RSH(l, -m)] = 0.0;
    }
    if (m == 0) {
      double v = 1.0;
      for (int l = 0; l <= lmax; l++) {
        if (l > 0 && -1.0 == x) v = -v;
        out[RSH(l, 0)] = v;
      }
    }
  } else {
    const double s1 = sqrt(1.0 - x * x);
    // P_mm, P_(m+1)m, then the recurrence
    double pmm;
    if (m == 0) {
      pmm = 1.0;
    } else {
      double sl = s1;
      for (int l = 2; l <= m; l++) sl = sl * s1;
      pmm = sl * dfac[2 * m - 1];
    }
    out[RSH(m, m)] = pmm;
    if (m + 1 <= lmax) out[RSH(m + 1, m)] = x * (2 * m + 1) * out[RSH(m, m)];
    for (int l = m + 2; l <= lmax; l++)
      out[RSH(l, m)] = (x * (2 * l - 1) * out[RSH(l - 1, m)] - (l + m - 1) * out[RSH(l - 2, m)]) / (l - m);
  }
  double sm = 0.0, cm = 1.0;
  for (int k = 1; k <= m; k++) {
    if (0.0 == phi) {
      sm = 0.0;
      cm = 1.0;
    } else if (k == 1) {
      sm = sp;
      cm = cp;
    } else {
      const double ns = sp * cm + cp * sm;
      cm = cp * cm - sp * sm;
      sm = ns;
    }
  }
  for (int l = m; l <= lmax; l++) {
    const double norm0 = sqrt((2.0 * l + 1.0) / (2.0 * M_PI));
    if (m == 0) {
      out[RSH(l, 0)] = norm0 * out[RSH(l, 0)];
      out[RSH(l, 0)] = out[RSH(l, 0)] / sqrt(2.0);
    } else {
      const double P = (sqrt(fac[l - m] / fac[l + m]) * norm0) * out[RSH(l, m)];
      out[RSH(l, -m)] = P * sm;
      out[RSH(l, m)] = P * cm;
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

// U tables of one (centre, row, chunk of 128 points): row C.L is the local
// channel U_L (type 1), rows l < L the semi-local channels times r^k,
// k < UK (type 2), zero past the channel's cut-off end[l].  The type-2 rows
// are Ut[l][k][n], or with J.ptmajor point-major, Ut[l][n][k]: the lanes of
// a type-2 group then read the same point of neighbouring k together.
__device__ void utab_rows(const DevTables &T, const DevJob &J, int c, int row, int chunk) {
  const ecp_centre C = J.cent[c];
  if (row > C.L) return;
  const int n = chunk * 128 + (int)threadIdx.x;
  if (n >= SO) return;
  const ecp_gauss *g = J.gauss + C.g0;
  const double r = T.xs[n];
  if (row == C.L) {
    J.UL[C.UL_off + n] = eval_ecp(g, C.ng, C.L, r);
    return;
  }
  const int l = row;
  double P = (n <= C.end[l]) ? eval_ecp(g, C.ng, l, r) : 0.0;
  const size_t rs = J.ptmajor ? 1 : SO;  // stride between the powers k
  double *dst = J.Ut + C.Ut_off + (size_t)l * J.UK * SO + (J.ptmajor ? (size_t)n * J.UK : (size_t)n);
  dst[0] = P;
  for (int k = 1; k < J.UK; k++) {
    P = P * r;
    dst[k * rs] = P;
  }
}

// F[centre][run][shell][lambda][n] (type2.c:263-303), or with J.ptmajor
// F[centre][run][shell][n][lambda] (the lanes of a type-2 group then read
// the same point of neighbouring lambda together), one block per
// non-skipped (centre, shell, run); the whole slab is written (zeros outside
// the window).
template <int LM>
__device__ void ftab_slab(const DevTables &T, const DevJob &J, const ecp_centre &C, int sl, int run,
                          const ecp_window &W) {
  const ecp_shell S = J.shells[C.shell0 + sl];
  const double dAC = dist(J.atoms[C.atom], J.atoms[S.atom]);
  double *slab = J.F + C.F_off + ((size_t)run * C.nsh + sl) * J.FL * SO;
  for (int n = threadIdx.x; n < SO; n += blockDim.x) {
    double acc[LM + 1];
#pragma unroll
    for (int i = 0; i <= LM; i++) acc[i] = 0.0;
    if (n >= W.start && n < W.end) {
      const double r = T.xs[n];
      for (int p = 0; p < S.K; p++) {
        const double zeta = J.pa[S.prim + p];
        double da = J.pd[S.prim + p];
        // derivatives: run n carries a^n (type2.c:286-289)
        for (int k = 0; k < run; k++) da *= zeta;
        double K[LM + 1];
        bessel_k<LM>(T.bessel, 2.0 * zeta * dAC * r, K);
        double e = dAC - r;
        e = exp(-zeta * e * e);
#pragma unroll
        for (int i = 0; i <= LM; i++) acc[i] += da * K[i] * e;
      }
    }
    const size_t rs = J.ptmajor ? 1 : SO;  // stride between the orders lambda
    double *row = slab + (J.ptmajor ? (size_t)n * J.FL : (size_t)n);
#pragma unroll
    for (int i = 0; i <= LM; i++) row[i * rs] = acc[i];
    for (int i = LM + 1; i < J.FL; i++) row[i * rs] = 0.0;
  }
}

__device__ void ftab_item(const DevTables &T, const DevJob &J, int2 it) {
  const int sl = it.y >> 2, run = it.y & 3;  // (shell, run) packed
  const ecp_centre C = J.cent[it.x];
  const ecp_window W = J.win[C.win_off + sl];
  const int lm = C.L - 1 + J.shells[C.shell0 + sl].l + run;
  switch (lm) {
#define FT(k) \
  case k:     \
    ftab_slab<k>(T, J, C, sl, run, W); \
    break;
    FT(0) FT(1) FT(2) FT(3) FT(4) FT(5) FT(6) FT(7) FT(8) FT(9) FT(10) FT(11) FT(12)
#undef FT
    default:
      break;
  }
}

// per-(centre, atom) angular table ang[lambda][LM(l,m)][C(la,c)]
constexpr int kAngRsph = (ECP_MAX_L + ECP_MAX_LBS) * (ECP_MAX_L + ECP_MAX_LBS);

__device__ void ang_item(const DevTables &T, const DevJob &J, int2 it, const double *__restrict__ fac,
                         const double *__restrict__ dfac, double *rsph) {
  const ecp_centre C = J.cent[it.x];
  const int A = C.atom0 + it.y;
  const int lmax_a = J.atom_maxl[A] + J.deriv;
  const int lmax = C.L - 1 + lmax_a;
  if (threadIdx.x == 0) {
    const ecp_xyz a = J.atoms[A], c = J.atoms[C.atom];
    double r, th, ph;
    sph_coords(a.x - c.x, a.y - c.y, a.z - c.z, r, th, ph);
    real_sph_inplace(lmax, th, ph, fac, dfac, rsph);
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
  // derivatives: coefficients times zeta^shift (type1.c:239-246)
  for (int k = 0; k < tr.powa; k++) G.ca *= G.za;
  for (int k = 0; k < tr.powb; k++) G.cb *= G.zb;
  for (int p = 0; p < 3; p++) G.P[p] = 2.0 * (G.za * rAC[p] + G.zb * rBC[p]);
  G.s = sqrt(G.P[0] * G.P[0] + G.P[1] * G.P[1] + G.P[2] * G.P[2]);
  G.zd2 = -G.za * G.dAC * G.dAC - G.zb * G.dBC * G.dBC;
  G.z = -G.za - G.zb;
}

// Results of a finished primitive pair: converged entries, zeros for the
// others, which go to the large-grid fallback queue.
template <int NQ>
__device__ __forceinline__ void t1_finish(int u, const T1Unit &U, const double (&p)[NQ], unsigned long long done,
                                          double *__restrict__ qout, unsigned long long *__restrict__ qfail,
                                          int *__restrict__ fbq, int *__restrict__ fbcount,
                                          unsigned char *__restrict__ fbflag) {
  const unsigned long long all = (NQ >= 64) ? ~0ull : ((1ull << NQ) - 1ull);
  double *out = qout + U.qoff;
#pragma unroll
  for (int e = 0; e < NQ; e++) out[e] = (done & (1ull << e)) ? p[e] : 0.0;
  const unsigned long long failed = all & ~done;
  qfail[u] = failed;
  if (failed) {
    fbq[atomicAdd(fbcount, 1)] = u;
    fbflag[U.trip] = 1;
  }
}

// Type-1 small grid, one lane per primitive pair (MODE 0: the whole PS93
// sequence; MODE 1, head: levels [0, head), then a pair still open saves its
// integrator state -- p, q, I per entry in planes of `plane` doubles at
// state -- and its converged-entry mask and is queued for a tail kernel;
// MODE 2, lane-serial tail: the queued pairs, compacted one per lane, resume
// at level `head`).  Within a level each point pair's sum
// T = (0 + f(left)) + f(right) is added to I in pair order, and the PS93
// recurrences and test (gc_integrators.c:156-217) follow.
template <int LAB, int MODE>
__device__ __forceinline__ void type1_body(int bid, int nb, const DevTables &T, const DevJob &J,
                                           const T1Unit *__restrict__ units, int nunits, double *__restrict__ qout,
                                           unsigned long long *__restrict__ qfail, int *__restrict__ fbq,
                                           int *__restrict__ fbcount, int head, double *__restrict__ state,
                                           size_t plane, int *__restrict__ tq, int *__restrict__ tcount,
                                           unsigned char *__restrict__ fbflag,
                                           unsigned long long *__restrict__ qopen = nullptr) {
  constexpr int NQ = nq_of(LAB);
  const int count = MODE == 2 ? *tcount : nunits;
  const int stride = MODE == 2 ? nb * 64 : count;
  for (int it = bid * 64 + (int)threadIdx.x; it < count; it += stride) {
    const int u = MODE == 2 ? tq[it] : it;
    const T1Unit U = units[u];
    const ecp_triplet tr = J.trip[U.trip];
    const ecp_centre Cn = J.cent[tr.centre];
    T1Geom G;
    t1_geom(J, tr, U.pa, U.pb, G);
    const double C0 = G.ca * G.cb * exp(G.zd2);
    const double *__restrict__ UL = J.UL + Cn.UL_off;
    const int gs = tr.gs, ge = tr.ge;
    const double minExp = T.lnAcc1;
    // w_n f(n) of one point for every entry, added into acc[] (type1.c:78-88
    // integrand, tabulated only on [gs, ge) as type1.c:121-130)
    auto f = [&](int n, double (&acc)[NQ]) {
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
    double p[NQ], q[NQ], I[NQ], Tt[NQ];
    unsigned long long done = 0ull;
    const unsigned long long all = (NQ >= 64) ? ~0ull : ((1ull << NQ) - 1ull);
    int lv0 = 2, lv1 = T.s93.nlevels;
    if (MODE == 2) {
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        p[e] = state[U.qoff + e];
        q[e] = state[plane + U.qoff + e];
        I[e] = state[2 * plane + U.qoff + e];
      }
      done = qfail[u];
      lv0 = head;
    } else {
#pragma unroll
      for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
      f(T.s93.centre, Tt);
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        p[e] = Tt[e];
        Tt[e] = 0.0;
      }
      f(T.s93.first_l, Tt);
      double t2[NQ];
#pragma unroll
      for (int e = 0; e < NQ; e++) t2[e] = 0.0;
      f(T.s93.first_r, t2);
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        q[e] = Tt[e] + t2[e];
        I[e] = p[e] + q[e];
      }
      if (MODE == 1) lv1 = min(head, T.s93.nlevels);
    }
    for (int lv = lv0; lv < lv1; lv++) {
      const int j = T.s93.lvl_j[lv], nn = T.s93.lvl_n[lv];
      int cnt = 0;
      for (int k = T.s93.lvl_begin[lv]; k < T.s93.lvl_begin[lv + 1]; k++) {
        const int li = T.p93l[k], ri = T.p93r[k];
#pragma unroll
        for (int e = 0; e < NQ; e++) Tt[e] = 0.0;
        if (li >= gs) {
          cnt++;
          f(li, Tt);
        }
        if (ri <= ge) {
          cnt++;
          f(ri, Tt);
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
    if (MODE == 1 && done != all && lv1 < T.s93.nlevels) {
#pragma unroll
      for (int e = 0; e < NQ; e++) {
        state[U.qoff + e] = p[e];
        state[plane + U.qoff + e] = q[e];
        state[2 * plane + U.qoff + e] = I[e];
      }
      qfail[u] = done;
      // the entries still open, for a fallback that starts beside the tail
      if (qopen) qopen[u] = all & ~done;
      tq[atomicAdd(tcount, 1)] = u;
      continue;
    }
    t1_finish<NQ>(u, U, p, done, qout, qfail, fbq, fbcount, fbflag);
  }
}

template <int LAB, int MODE>
__global__ __launch_bounds__(64) void k_type1(DevTables T, DevJob J, const T1Unit *__restrict__ units, int nunits,
                                              double *__restrict__ qout, unsigned long long *__restrict__ qfail,
                                              int *__restrict__ fbq, int *__restrict__ fbcount, int head,
                                              double *__restrict__ state, size_t plane, int *__restrict__ tq,
                                              int *__restrict__ tcount, unsigned char *__restrict__ fbflag) {
  type1_body<LAB, MODE>(blockIdx.x, gridDim.x, T, J, units, nunits, qout, qfail, fbq, fbcount, head, state, plane,
                        tq, tcount, fbflag);
}

// f + a[0] b[0] + a[sa] b[sb] + ... (n terms), accumulated in order; the
// reads of up to B terms are issued before their products are added
template <int B>
__device__ __forceinline__ double dot_in_order(double f, const double *a, int sa, const double *b, int sb, int n) {
  int m = 0;
  for (; m + B <= n; m += B) {
    double x[B], y[B];
#pragma unroll
    for (int i = 0; i < B; i++) {
      x[i] = a[(m + i) * sa];
      y[i] = b[(m + i) * sb];
    }
#pragma unroll
    for (int i = 0; i < B; i++) f += x[i] * y[i];
  }
  for (; m < n; m++) f += a[m * sa] * b[m * sb];
  return f;
}

// Wave-level helpers (blocks of exactly one 64-lane wave).
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// I + v[0] + v[1] + ... + v[n - 1], added left to right (the order of a
// lane-serial integrator); the reads (LDS or global) are issued eight
// (then four) at a time so that one wait covers several of them
__device__ __forceinline__ double sum_in_order(double I, const double *v, int n) {
  int k = 0;
  for (; k + 8 <= n; k += 8) {
    double x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = v[k + i];
#pragma unroll
    for (int i = 0; i < 8; i++) I += x[i];
  }
  if (k + 4 <= n) {
    double x[4];
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = v[k + i];
#pragma unroll
    for (int i = 0; i < 4; i++) I += x[i];
    k += 4;
  }
  for (; k < n; k++) I += v[k];
  return I;
}

// Tail: one wave per primitive pair left open by the head, levels [head,
// nlevels).  Lane k evaluates the k-th point pair of a round of levels (at
// most 64 pairs per round; a level of the 383-point grid has at most 64) for
// every Q entry and stores the pair sum T_k = (0 + f(left)) + f(right) in
// LDS; lane e then adds each level's T_k
// to its entry's I in pair order and applies the PS93 test, so the
// arithmetic is that of the lane-serial integrator.
constexpr int kTS = 65;  // LDS stride of the per-entry pair sums (bank spread)

template <int LAB>
__device__ __forceinline__ void t1tail_body(int bid, int nb, const DevTables &T, const DevJob &J,
                                            const T1Unit *__restrict__ units, double *__restrict__ qout,
                                            unsigned long long *__restrict__ qfail, int *__restrict__ fbq,
                                            int *__restrict__ fbcount, int head, const double *__restrict__ state,
                                            size_t plane, const int *__restrict__ tq, const int *__restrict__ tcount,
                                            unsigned char *__restrict__ fbflag, double *s_T, bool spec = false) {
  constexpr int NQ = nq_of(LAB);
  static_assert(NQ <= 64, "one lane per Q entry");
  const int count = *tcount;
  const int lane = threadIdx.x;
  for (int it = bid; it < count; it += nb) {
    const int u = tq[it];
    const T1Unit U = units[u];
    const ecp_triplet tr = J.trip[U.trip];
    const ecp_centre Cn = J.cent[tr.centre];
    T1Geom G;
    t1_geom(J, tr, U.pa, U.pb, G);
    const double C0 = G.ca * G.cb * exp(G.zd2);
    const double *__restrict__ UL = J.UL + Cn.UL_off;
    const int gs = tr.gs, ge = tr.ge;
    const double minExp = T.lnAcc1;
    auto f = [&](int n, double (&acc)[NQ]) {
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
    const unsigned long long done0 = qfail[u];
    double p = 0.0, q = 0.0, I = 0.0;
    bool mydone = true;
    if (lane < NQ) {
      p = state[U.qoff + lane];
      q = state[plane + U.qoff + lane];
      I = state[2 * plane + U.qoff + lane];
      mydone = (done0 >> lane) & 1ull;
    }
    // rounds of consecutive levels whose point pairs fit the wave together
    // (levels 8-11: 48 pairs, 12-13: 64, 14: 64); the sums and tests then
    // follow level by level, so only the evaluation is speculative
    int lv = head;
    while (lv < T.s93.nlevels) {
      const int b0 = T.s93.lvl_begin[lv];
      int l1 = lv + 1;
      while (l1 < T.s93.nlevels && T.s93.lvl_begin[l1 + 1] - b0 <= 64) l1++;
      const int tot = T.s93.lvl_begin[l1] - b0;
      int cnt = 0;
      if (lane < tot) {
        const int li = T.p93l[b0 + lane], ri = T.p93r[b0 + lane];
        double t[NQ];
#pragma unroll
        for (int e = 0; e < NQ; e++) t[e] = 0.0;
        if (li >= gs) {
          cnt++;
          f(li, t);
        }
        if (ri <= ge) {
          cnt++;
          f(ri, t);
        }
#pragma unroll
        for (int e = 0; e < NQ; e++) s_T[e * kTS + lane] = t[e];
      }
      const unsigned long long cmask = __ballot(cnt > 0);
      __syncthreads();
      bool stop = false;
      for (int l = lv; l < l1; l++) {
        const int o = T.s93.lvl_begin[l] - b0, np = T.s93.lvl_begin[l + 1] - T.s93.lvl_begin[l];
        const unsigned long long lm = (np >= 64) ? ~0ull : (((1ull << np) - 1ull) << o);
        const bool any = (cmask & lm) != 0ull;
        if (lane < NQ) I = sum_in_order(I, s_T + lane * kTS + o, np);
        if (!mydone) {
          const int j = T.s93.lvl_j[l], nn = T.s93.lvl_n[l];
          double err;
          if (j == 0) {
            p += (I - q);
            err = 16 * fabs(q - 3 * p / 2) / (3 * nn);
          } else {
            err = 16 * fabs(I - 2 * q) / (3 * nn);
            q = I;
          }
          if (any && err < T.tol) {
            mydone = true;
            p = 16 * q / (3 * nn);
          }
        }
        if (!wave_any(!mydone)) {
          stop = true;
          break;
        }
      }
      __syncthreads();
      if (stop) break;
      lv = l1;
    }
    if (lane < NQ) qout[U.qoff + lane] = mydone ? p : 0.0;
    const unsigned long long failed = __ballot(lane < NQ && !mydone);
    if (lane == 0) {
      qfail[u] = failed;
      if (failed) {
        // spec: the fallback already runs over every unit of the tail
        if (!spec) fbq[atomicAdd(fbcount, 1)] = u;
        fbflag[U.trip] = 1;
      }
    }
  }
}

template <int LAB>
__global__ __launch_bounds__(64) void k_t1tail(DevTables T, DevJob J, const T1Unit *__restrict__ units,
                                               double *__restrict__ qout, unsigned long long *__restrict__ qfail,
                                               int *__restrict__ fbq, int *__restrict__ fbcount, int head,
                                               const double *__restrict__ state, size_t plane,
                                               const int *__restrict__ tq, const int *__restrict__ tcount,
                                               unsigned char *__restrict__ fbflag) {
  __shared__ double s_T[nq_of(LAB) * kTS];
  t1tail_body<LAB>(blockIdx.x, gridDim.x, T, J, units, qout, qfail, fbq, fbcount, head, state, plane, tq, tcount,
                   fbflag, s_T);
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

// First large-grid index n with r_n > max(dAC, dBC) and the exponent below
// the numerical zero: the screening break of type1.c:222-225 /
// type2.c:479-481; `order` if there is none.  expo(r) is evaluated by the
// caller's functor with the same arithmetic as the integrand.
// The predicate "r_n > max(dAC, dBC) and expo(r_n) < minExp" is monotone in
// n (r_n increases with n, and beyond both centres the exponent decreases;
// every rounding step is monotone), so its first true index is found with
// two wave-wide probes instead of a scan: 64 samples at stride
// ceil(order / 64) bracket it, then the 64 points ending at the first true
// sample locate it.
template <typename Expo>
__device__ int fm06_break(const DevTables &T, const Fm06 &M, double dAC, double dBC, double minExp, Expo expo) {
  const int order = T.lorder;
  auto pred = [&](int n) {
    const double r = M.i1 * T.xl[n] + M.i2;
    return r > dAC && r > dBC && expo(r) < minExp;
  };
  const int stride = (order + 63) / 64;
  const int ns = min(order - 1, ((int)threadIdx.x + 1) * stride - 1);
  const unsigned long long b1 = __ballot(pred(ns));
  if (!b1) return order;
  const int k0 = __ffsll((long long)b1) - 1;
  const int hi = min(order - 1, (k0 + 1) * stride - 1);
  const int n = hi - 63 + (int)threadIdx.x;
  const unsigned long long b2 = __ballot(n >= 0 && pred(n));
  return hi - 63 + (__ffsll((long long)b2) - 1);
}

// PSM92 evaluation rounds: the points of consecutive levels are evaluated by
// the wave together when they fit (levels 0-1: 3 points, 2-5: 60, 6: 64);
// larger levels go in chunks of 32 point pairs.  Within a round the centre
// (level 0) takes slot 0 and pair k of level lv the slots roff + 2k (left)
// and roff + 2k + 1 (right), roff being the points of the round's earlier
// levels.  Summation order inside each level is unchanged.
__device__ __forceinline__ int psm92_points(int lv) { return lv == 0 ? 1 : (1 << lv); }

// hw: the slots of a round (lanes per integration, 16, 32 or 64); levels
// 2 .. l2 - 1 take 2^l2 - 4 <= hw of them
__device__ __forceinline__ int psm92_round_end(int lv, int nlev, int hw = 64) {
  const int l2 = hw >= 64 ? 6 : hw >= 32 ? 5 : 4;
  return min(nlev, lv == 0 ? 2 : (lv == 2 ? l2 : lv + 1));
}

// grid index evaluated by `slot` in the round [l0, l1); -1 for none (a slot
// past the round, or a right point past the screening end)
__device__ __forceinline__ int psm92_round_point(const DevTables &T, int l0, int l1, int slot, int end) {
  int o = 0;
  for (int lv = l0; lv < l1; lv++) {
    const int np = psm92_points(lv);
    if (slot < o + np) {
      if (lv == 0) return T.s92.centre;
      const int pair = T.s92.lvl_begin[lv] + ((slot - o) >> 1);
      if (((slot - o) & 1) == 0) return T.p92l[pair];
      const int ri = T.p92r[pair];
      return ri <= end ? ri : -1;
    }
    o += np;
  }
  return -1;
}

// type-1 fallback (type1.c:205-252): one wave per primitive pair with failed
// Q entries.  Per level the wave evaluates the level's points of the
// FM06-mapped PSM92 grid (Bessel functions, U_L, exponential) into LDS, then
// each lane owning a failed entry sums them in the integrator's pair order
// (gc_integrators.c:55-84) and applies the PSM92 test; the wave stops when
// every entry has converged.  Points past the screening break contribute 0.
constexpr int kKS = 65;     // LDS stride of per-point Bessel rows (bank spread)
constexpr int kFb1TS = 33;  // LDS stride of a failed entry's pair sums (at most 32 pairs per call)

template <int LAB, bool SPEC = false>
__device__ __forceinline__ void fb1_body(int bid, int nb, const DevTables &T, const DevJob &J,
                                         const T1Unit *__restrict__ units, double *__restrict__ qout,
                                         const unsigned long long *__restrict__ qfail, const int *__restrict__ fbq,
                                         const int *__restrict__ fbcount, unsigned char *__restrict__ status,
                                         double *s_r, double *s_w, double *s_u, double *s_g, double *s_k,
                                         double *s_t, int *s_e, const int *__restrict__ tq = nullptr,
                                         const int *__restrict__ tcount = nullptr,
                                         const unsigned long long *__restrict__ qopen = nullptr,
                                         double *__restrict__ qfb = nullptr,
                                         unsigned long long *__restrict__ qfbf = nullptr) {
  // qfb != nullptr (speculative): besides the units whose small-grid
  // integrals failed (fbq), every unit the tail takes (tq), for the entries
  // open after the head (qopen), started before the tail has decided them;
  // results to qfb, large-grid failures as an entry mask to qfbf, and the
  // assembly uses a result where the small grid failed
  const int n1 = *fbcount;
  int count = n1;
  if constexpr (SPEC) count += *tcount;
  const int lane = threadIdx.x;
  for (int it = bid; it < count; it += nb) {
    int u;
    unsigned long long mask;
    if (!SPEC || it < n1) {
      u = fbq[it];
      mask = qfail[u];
    } else {
      u = tq[it - n1];
      mask = qopen[u];
    }
    const T1Unit U = units[u];
    const ecp_triplet tr = J.trip[U.trip];
    const ecp_centre Cn = J.cent[tr.centre];
    T1Geom G;
    t1_geom(J, tr, U.pa, U.pb, G);
    const double Cc = G.ca * G.cb;
    const double zeta_p = G.za + G.zb;
    const Fm06 M = fm06(zeta_p, (G.za * G.dAC + G.zb * G.dBC) / zeta_p);
    const double minExp = T.lnAcc1;
    auto expo_of = [&](double r) { return (G.z * r + G.s) * r + G.zd2; };
    const int brk = fm06_break(T, M, G.dAC, G.dBC, minExp, expo_of);
    const int end = brk - 1;
    // this lane's failed entry (the lane-th set bit of mask)
    int my_e = -1, my_l1 = 0, my_l2 = 0;
    const int nfail = __popcll(mask);
    {
      int e = 0, seen = 0;
      for (int l1 = 0; l1 <= LAB; l1++)
        for (int l2 = l1; l2 >= 0; l2 -= 2, e++)
          if ((mask >> e) & 1ull) {
            if (seen == lane) {
              my_e = e;
              my_l1 = l1;
              my_l2 = l2;
            }
            seen++;
          }
    }
    const ecp_gauss *gs = J.gauss + Cn.g0;
    auto eval = [&](int slot, int n) {
      double r = 0.0, w = 0.0, uu = 0.0, g = 0.0;
      double K[LAB + 1];
#pragma unroll
      for (int l = 0; l <= LAB; l++) K[l] = 0.0;
      if (n >= 0 && n < brk) {
        r = M.i1 * T.xl[n] + M.i2;
        const double expo = expo_of(r);
        if (expo >= minExp) {
          g = exp(expo);
          bessel_k<LAB>(T.bessel, G.s * r, K);
          uu = eval_ecp(gs, Cn.ng, Cn.L, r);
          w = T.wl[n] * M.i1;
        }
      }
      s_r[slot] = r;
      s_w[slot] = w;
      s_u[slot] = uu;
      s_g[slot] = g;
#pragma unroll
      for (int l = 0; l <= LAB; l++) s_k[l * kKS + slot] = K[l];
    };
    auto fj = [&](int l1, int l2, int slot) {
      double rn = 1.0;
      for (int k = 0; k < l1; k++) rn = s_r[slot] * rn;
      return s_w[slot] * ((((Cc * rn) * s_u[slot]) * s_k[l2 * kKS + slot]) * s_g[slot]);
    };
    auto f = [&](int slot) { return fj(my_l1, my_l2, slot); };
    const bool mine = my_e >= 0;
    if (mine) s_e[lane] = my_l1 | (my_l2 << 8);
    double I = 0.0, p = 0.0, res = 0.0;
    bool done = !mine;
    auto test = [&](int l, double q) {
      const double N = T.s92.lvl_n[l];
      const double ee = I - p;
      if (16 * ee * ee <= 3 * N * fabs(I - q) * T.tol) {
        done = true;
        res = 16 * I / (3 * N);
      }
    };
    // pairs [kb, ke) of level l, pair k at slot0 + 2 (k - kb): the pair sums
    // (0 + f(left)) + f(right) of every open entry by all lanes into s_t,
    // then each entry's lane adds its own in pair order.  Called by the
    // whole wave.
    auto sum = [&](int l, int slot0, int kb, int ke) {
      const int b0 = T.s92.lvl_begin[l], np = ke - kb;
      const unsigned long long open = __ballot(mine && !done);
      for (int i = lane; i < nfail * np; i += 64) {
        const int j = i / np, k = i - j * np;
        if (!((open >> j) & 1ull)) continue;
        const int code = s_e[j];
        double t = fj(code & 0xff, code >> 8, slot0 + 2 * k);
        if (T.p92r[b0 + kb + k] <= end) t += fj(code & 0xff, code >> 8, slot0 + 2 * k + 1);
        s_t[j * kFb1TS + k] = t;
      }
      __syncthreads();
      if (mine && !done) I = sum_in_order(I, s_t + lane * kFb1TS, np);
      __syncthreads();
    };
    int lv = 0;
    while (lv < T.s92.nlevels && wave_any(!done)) {
      const int l1 = psm92_round_end(lv, T.s92.nlevels);
      if (l1 > lv + 1 || psm92_points(lv) <= 64) {
        eval(lane, psm92_round_point(T, lv, l1, lane, end));
        __syncthreads();
        int roff = 0;
        for (int l = lv; l < l1; l++) {
          if (l == 0) {
            if (mine) I = p = f(0);
          } else {
            double q = 0.0;
            if (!done) {
              q = 2 * p;
              p = 2 * I;
            }
            sum(l, roff, 0, T.s92.lvl_begin[l + 1] - T.s92.lvl_begin[l]);
            if (!done) test(l, q);
          }
          roff += psm92_points(l);
        }
        __syncthreads();
        lv = l1;
      } else {
        double q = 0.0;
        if (!done) {
          q = 2 * p;
          p = 2 * I;
        }
        const int b0 = T.s92.lvl_begin[lv], b1 = T.s92.lvl_begin[lv + 1];
        for (int c0 = b0; c0 < b1; c0 += 32) {
          const int nc = min(32, b1 - c0);
          const int k = lane >> 1;
          int n = -1;
          if (k < nc) {
            if ((lane & 1) == 0) n = T.p92l[c0 + k];
            else if (T.p92r[c0 + k] <= end) n = T.p92r[c0 + k];
          }
          eval(lane, n);
          __syncthreads();
          sum(lv, 0, c0 - b0, c0 - b0 + nc);
        }
        if (!done) test(lv, q);
        lv++;
      }
    }
    if constexpr (SPEC) {
      if (mine && done) qfb[U.qoff + my_e] = res;
      // the failed entries as a mask: one bit per lane that owns one
      unsigned int lo = (mine && !done && my_e < 32) ? (1u << my_e) : 0u;
      unsigned int hi = (mine && !done && my_e >= 32) ? (1u << (my_e - 32)) : 0u;
      for (int o = 32; o > 0; o >>= 1) {
        lo |= (unsigned int)__shfl_xor((int)lo, o, 64);
        hi |= (unsigned int)__shfl_xor((int)hi, o, 64);
      }
      if (lane == 0) qfbf[u] = ((unsigned long long)hi << 32) | lo;
    } else if (mine) {
      if (done) qout[U.qoff + my_e] = res;
      else status[2 * U.trip] = 1;
    }
    __syncthreads();
  }
}

template <int LAB>
__global__ __launch_bounds__(64) void k_fb1(DevTables T, DevJob J, const T1Unit *__restrict__ units,
                                            double *__restrict__ qout, const unsigned long long *__restrict__ qfail,
                                            const int *__restrict__ fbq, const int *__restrict__ fbcount,
                                            unsigned char *__restrict__ status) {
  __shared__ double s_r[64], s_w[64], s_u[64], s_g[64], s_k[(LAB + 1) * kKS], s_t[nq_of(LAB) * kFb1TS];
  __shared__ int s_e[64];
  fb1_body<LAB>(blockIdx.x, gridDim.x, T, J, units, qout, qfail, fbq, fbcount, status, s_r, s_w, s_u, s_g, s_k,
                s_t, s_e);
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

// One T entry: PS93 up to level lv_end (exclusive).  Returns 0 when it
// converged (result in *res), 1 when every level ran without convergence,
// 2 when it stopped unconverged at lv_end.
__device__ __forceinline__ int t2_entry(const DevTables &T, const DevJob &J, const T2Group &G, int code, int lv_end,
                                        double *res) {
  const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
  const ecp_triplet tr = J.trip[G.trip];
  const ecp_centre C = J.cent[tr.centre];
  // point and row strides of the F and U rows (layouts as the tables kernel writes them)
  const int pF = J.ptmajor ? J.FL : 1, pU = J.ptmajor ? J.UK : 1, rs = J.ptmajor ? 1 : SO;
  const double *__restrict__ Fa =
      J.F + C.F_off + ((size_t)tr.powa * C.nsh + (tr.sa - C.shell0)) * J.FL * SO + (size_t)l1 * rs;
  const double *__restrict__ Fb =
      J.F + C.F_off + ((size_t)tr.powb * C.nsh + (tr.sb - C.shell0)) * J.FL * SO + (size_t)l2 * rs;
  const double *__restrict__ Uu = J.Ut + C.Ut_off + (size_t)G.l * J.UK * SO + (size_t)l3 * rs;
  const int gs = tr.gs, ge = tr.ge;
  auto f = [&](int n) { return T.ws[n] * ((Fa[n * pF] * Fb[n * pF]) * Uu[n * pU]); };
  double p = f(T.s93.centre);
  double q = f(T.s93.first_l) + f(T.s93.first_r);
  double I = p + q;
  *res = 0.0;
  for (int lv = 2; lv < lv_end; lv++) {
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
      *res = 16 * q / (3 * nn);
      return 0;
    }
  }
  return lv_end < T.s93.nlevels ? 2 : 1;
}

struct T2Args {
  const T2Group *groups;
  const int *goff;
  int ngroups, ntotal;
  const int *tmpl;
  double *Tout;
  unsigned char *tfail;
  int *gflag, *fbq, *fbcount;  // fbq: per class, the triplets with failed entries (qstride apart)
  int qstride;
  unsigned char *fbflag;
  int *flist;
  int *tq, *tcount;  // entries left open by the head
  int *tclaim;                  // per triplet: queued for the fallback
  const unsigned char *tcls2;   // per triplet: its fallback class max(la, lb) + L - 1
  int spec;                     // the fallback takes every entry the head leaves open
  unsigned char *fbf2;          // spec: per entry, the fallback failed on the large grid
};

// An entry for the large-grid fallback: gflag[g] counts the group's
// entries, listed (in any order) at flist[eoff ...]; the triplet's first one
// queues the triplet, whose listed entries of all groups the fallback takes
// together
__device__ __forceinline__ void t2_enlist(const T2Args &A, int gid, int g, const T2Group &G) {
  // the speculative fallback's failure flag of the entry, read only for
  // enlisted entries: cleared here, before the fallback runs
  if (A.spec) A.fbf2[gid] = 0;
  const int k = atomicAdd(&A.gflag[g], 1);
  A.flist[G.eoff + k] = gid - G.eoff;
  if (k == 0 && atomicExch(&A.tclaim[G.trip], 1) == 0) {
    const int c = A.tcls2[G.trip];
    const int slot = atomicAdd(&A.fbcount[c], 1);
    A.fbq[(size_t)c * A.qstride + slot] = G.trip;
    A.fbflag[G.trip] = 1;
  }
}

// A decided entry: its small-grid result and failure flag; a failed entry
// goes to the fallback, unless the fallback already took it speculatively
__device__ __forceinline__ void t2_store(const T2Args &A, int gid, int g, const T2Group &G, int st, double res) {
  A.Tout[gid] = res;
  A.tfail[gid] = st == 0 ? 0 : 1;
  if (st != 0 && !A.spec) t2_enlist(A, gid, g, G);
}

// One lane per T entry.  With head < the level count, entries still open
// after `head` levels are queued instead for k_type2_tail, so that the few
// stragglers neither hold a wave of converged lanes nor run their remaining
// levels one point at a time.
__global__ __launch_bounds__(128) void k_type2(DevTables T, DevJob J, T2Args A, int head) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= A.ntotal) return;
  const int g = find_group(A.goff, A.ngroups, gid);
  const T2Group G = A.groups[g];
  const int code = A.tmpl[G.tmpl + (gid - G.eoff)];
  double res;
  const int st = t2_entry(T, J, G, code, min(head, T.s93.nlevels), &res);
  if (st == 2) {
    A.tq[atomicAdd(A.tcount, 1)] = gid;
    if (A.spec) t2_enlist(A, gid, g, G);
    return;
  }
  t2_store(A, gid, g, G, st, res);
}

// Tail: the entries left open by the head, rerun whole by kT2Group lanes
// each (64 / kT2Group entries per wave).  Per level the group's lanes form
// the point-pair sums T_k = (0 + f(left)) + f(right) into LDS and the
// group's first lane adds them to I in pair order and applies the PS93 test,
// so the arithmetic is that of the lane-serial integrator.
constexpr int kT2Group = 4;

__global__ __launch_bounds__(64) void k_type2_tail(DevTables T, DevJob J, T2Args A) {
  constexpr int NE = 64 / kT2Group;
  __shared__ double s_t[NE][64];
  __shared__ int s_done[NE];
  const int lane = threadIdx.x, e = lane / kT2Group, r = lane % kT2Group;
  const int count = *A.tcount;
  for (int it0 = blockIdx.x * NE; it0 < count; it0 += gridDim.x * NE) {
    const int it = it0 + e;
    const bool live = it < count;
    int gid = 0, g = 0, code = 0;
    T2Group G{};
    if (live) {
      gid = A.tq[it];
      g = find_group(A.goff, A.ngroups, gid);
      G = A.groups[g];
      code = A.tmpl[G.tmpl + (gid - G.eoff)];
    }
    const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
    const ecp_triplet tr = J.trip[G.trip];
    const ecp_centre C = J.cent[tr.centre];
    // point and row strides of the F and U rows (layouts as the tables kernel writes them)
    const int pF = J.ptmajor ? J.FL : 1, pU = J.ptmajor ? J.UK : 1, rs = J.ptmajor ? 1 : SO;
    const double *__restrict__ Fa =
        J.F + C.F_off + ((size_t)tr.powa * C.nsh + (tr.sa - C.shell0)) * J.FL * SO + (size_t)l1 * rs;
    const double *__restrict__ Fb =
        J.F + C.F_off + ((size_t)tr.powb * C.nsh + (tr.sb - C.shell0)) * J.FL * SO + (size_t)l2 * rs;
    const double *__restrict__ Uu = J.Ut + C.Ut_off + (size_t)G.l * J.UK * SO + (size_t)l3 * rs;
    const int gs = tr.gs, ge = tr.ge;
    auto f = [&](int n) { return T.ws[n] * ((Fa[n * pF] * Fb[n * pF]) * Uu[n * pU]); };
    double p = 0.0, q = 0.0, I = 0.0, res = 0.0;
    bool done = !live;
    if (live && r == 0) {
      p = f(T.s93.centre);
      q = f(T.s93.first_l) + f(T.s93.first_r);
      I = p + q;
    }
    int st = 1;
    for (int lv = 2; lv < T.s93.nlevels; lv++) {
      if (!wave_any(!done)) break;
      const int b0 = T.s93.lvl_begin[lv], np = T.s93.lvl_begin[lv + 1] - b0;
      if (!done)
        for (int k = r; k < np; k += kT2Group) {
          const int li = T.p93l[b0 + k], ri = T.p93r[b0 + k];
          double t = 0.0;
          if (li >= gs) t += f(li);
          if (ri <= ge) t += f(ri);
          s_t[e][k] = t;
        }
      __syncthreads();
      if (!done && r == 0) {
        int cnt = 0;
        for (int k = 0; k < np; k++) {
          cnt += (T.p93l[b0 + k] >= gs) + (T.p93r[b0 + k] <= ge);
          I += s_t[e][k];
        }
        const int j = T.s93.lvl_j[lv], nn = T.s93.lvl_n[lv];
        double err;
        if (j == 0) {
          p += (I - q);
          err = 16 * fabs(q - 3 * p / 2) / (3 * nn);
        } else {
          err = 16 * fabs(I - 2 * q) / (3 * nn);
          q = I;
        }
        if (cnt > 0 && err < T.tol) {
          res = 16 * q / (3 * nn);
          st = 0;
          done = true;
        }
        s_done[e] = done;
      }
      __syncthreads();
      done = s_done[e] || !live;
      __syncthreads();
    }
    if (live && r == 0) t2_store(A, gid, g, G, st, res);
  }
}

// type-2 fallback (calcT_FM06, type2.c:417-528): one wave per (triplet with
// failed T entries, primitive pair).  The FM06-mapped points depend on the
// pair only, so the wave evaluates each level's points once (K_a, K_b, the
// exponential, and U_l for every channel l with failed entries) into LDS
// for all failed entries of all the triplet's channels; lanes own the failed
// entries (64 per pass) and sum the points in the integrator's pair order.
// Results per (entry, pair) go to fbT and are summed over pairs, in pair
// order, by the assembly.// type-2 fallback (calcT_FM06, type2.c:417-528): one wave per (triplet with
// failed T entries, primitive pair).  The FM06-mapped points depend on the
// pair only, so the wave evaluates each level's points once (K_a, K_b, the
// exponential, and U_l for every channel l with failed entries) into LDS
// for all failed entries of all the triplet's channels; lanes own the
// failed entries (64 per pass) and sum the points in the integrator's pair
// With PACK, a triplet with at most 16 (32) failed entries runs four (two)
// of its pairs per wave instead, one per lane group of HW = 16 (32) lanes,
// so that the lanes summing are not a few of 64: the item of every fourth
// (second) pair takes them, the others none; the groups share the wave's
// rounds, each over its own pair's points.
constexpr int kFb2PerLane = 1;
constexpr int kFb2Cnt = 52;  // fbc[kFb2Cnt + c]: triplets queued for the type-2 fallback, class c
constexpr int kFbcSlots = 80;  // queue counters per job
static_assert(ECP_MAX_LAB < 16 && kFb2Cnt + kMaxLMB < kFbcSlots, "queue counter slots");

template <int LMB, bool PACK>
__device__ __forceinline__ void fb2_body(int bid, int nb, const DevTables &T, const DevJob &J,
                                         const T2Group *__restrict__ groups, const int *__restrict__ tmpl,
                                         const unsigned char *__restrict__ tfail, const int *__restrict__ fbq,
                                         const int *__restrict__ fbcount, double *__restrict__ fbT,
                                         unsigned char *__restrict__ status, const int *__restrict__ gcount,
                                         const int *__restrict__ flist, const int *__restrict__ tgfirst,
                                         int nsplit, double *s_r, double *s_w, double *s_cu, double *s_ex,
                                         double *s_ka, double *s_kb, unsigned char *__restrict__ fbf2) {
  const int lane = threadIdx.x;
  // items (split, triplet, pair), split-major so that the grid-stride loop
  // spreads every split over all workgroups; split k takes the passes k,
  // k + nsplit, ...
  const long per_split = (long)(*fbcount) * J.maxPairs;
  const long total = per_split * nsplit;
  for (long it = bid; it < total; it += nb) {
    const int split = (int)(it / per_split);
    const long r = it - split * per_split;
    const int t = fbq[r / J.maxPairs];
    const int pair = (int)(r % J.maxPairs);
    const ecp_triplet tr = J.trip[t];
    const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
    const int npairs = Sa.K * Sb.K;
    if (pair >= npairs) continue;
    const ecp_centre Cn = J.cent[tr.centre];
    // the triplet's groups (one per semi-local channel l) and their failed
    // entries; the points are shared by all of them
    const int g0 = tgfirst[t], L = Cn.L;
    int gc[ECP_MAX_L], nfail = 0, lmax = 0;
    for (int l = 0; l < L; l++) {
      gc[l] = gcount[g0 + l];
      nfail += gc[l];
      if (gc[l]) lmax = l;
    }
    // this wave's passes: split, split + nsplit, ... (none: nothing to do)
    if (split * 64 * kFb2PerLane >= nfail) continue;
    // lane groups of HW = 1 << hwl lanes, group g on pair + g (a group past
    // the triplet's pairs idles on pair's points)
    const int hwl = !PACK ? 6 : nfail <= 16 ? 4 : nfail <= 32 ? 5 : 6;
    if (PACK && (pair & ((64 >> hwl) - 1))) continue;
    const int HW = PACK ? 1 << hwl : 64, NG = PACK ? 64 >> hwl : 1;
    const int grp = PACK ? lane >> hwl : 0, sl = PACK ? lane & (HW - 1) : lane, base = PACK ? grp << hwl : 0;
    const int mypair = pair + grp;
    const bool gvalid = !PACK || mypair < npairs;
    const int pq = gvalid ? mypair : pair;
    const int pa = pq / Sb.K, pb = pq % Sb.K;
    const ecp_xyz rc = J.atoms[Cn.atom];
    const double dAC = dist(rc, J.atoms[Sa.atom]), dBC = dist(rc, J.atoms[Sb.atom]);
    const double za = J.pa[Sa.prim + pa], zb = J.pa[Sb.prim + pb];
    const double s1 = 2.0 * za * dAC, s2 = 2.0 * zb * dBC;
    double Cc = J.pd[Sa.prim + pa] * J.pd[Sb.prim + pb];
    // derivatives: times zeta^shift (type2.c:457-462)
    for (int k = 0; k < tr.powa; k++) Cc *= za;
    for (int k = 0; k < tr.powb; k++) Cc *= zb;
    const double zeta_p = za + zb;
    const Fm06 M = fm06(zeta_p, (za * dAC + zb * dBC) / zeta_p);
    const double minExp = T.lnAcc2;
    auto expo_of = [&](double r) {
      const double d1 = dAC - r, d2 = dBC - r;
      return -za * d1 * d1 - zb * d2 * d2;
    };
    // the search is wave-wide: with lane groups on different pairs, the
    // whole wave finds each group's break in turn
    int brk = 0;
    if (NG == 1) {
      brk = fm06_break(T, M, dAC, dBC, minExp, expo_of);
    } else {
      for (int q = 0; q < NG; q++) {
        const int pq2 = pair + q < npairs ? pair + q : pair;
        const double za2 = J.pa[Sa.prim + pq2 / Sb.K], zb2 = J.pa[Sb.prim + pq2 % Sb.K];
        const double zp2 = za2 + zb2;
        const Fm06 M2 = fm06(zp2, (za2 * dAC + zb2 * dBC) / zp2);
        const int b2 = fm06_break(T, M2, dAC, dBC, minExp, [&](double r) {
          const double d1 = dAC - r, d2 = dBC - r;
          return -za2 * d1 * d1 - zb2 * d2 * d2;
        });
        if (q == grp) brk = b2;
      }
    }
    const int end = brk - 1;
    const int laC = tr.la + lmax, lbC = tr.lb + lmax;
    const ecp_gauss *gg = J.gauss + Cn.g0;
    auto eval = [&](int slot, int n) {
      double r = 0.0, w = 0.0, cu = 0.0, ex = 0.0;
      bool live = false;
      if (n >= 0 && n < brk) {
        r = M.i1 * T.xl[n] + M.i2;
        const double expo = expo_of(r);
        if (expo >= minExp) {
          live = true;
          ex = exp(expo);
          w = T.wl[n] * M.i1;
        }
      }
      for (int l = 0; l < L; l++) {
        if (!gc[l]) continue;
        cu = live ? Cc * eval_ecp(gg, Cn.ng, l, r) : 0.0;
        s_cu[l * 64 + slot] = cu;
      }
      s_r[slot] = r;
      s_w[slot] = w;
      s_ex[slot] = ex;
      // one Bessel vector live at a time (register pressure)
      {
        double K[LMB + 1];
#pragma unroll
        for (int l = 0; l <= LMB; l++) K[l] = 0.0;
        if (live) bessel_k<LMB>(T.bessel, s1 * r, K);
#pragma unroll
        for (int l = 0; l <= LMB; l++)
          if (l <= laC) s_ka[l * kKS + slot] = K[l];
      }
      {
        double K[LMB + 1];
#pragma unroll
        for (int l = 0; l <= LMB; l++) K[l] = 0.0;
        if (live) bessel_k<LMB>(T.bessel, s2 * r, K);
#pragma unroll
        for (int l = 0; l <= LMB; l++)
          if (l <= lbC) s_kb[l * kKS + slot] = K[l];
      }
    };
    // the failed entries of all groups (listed by k_type2), group after
    // group, HW per pass
    for (int pass0 = split * HW * kFb2PerLane; pass0 < nfail; pass0 += HW * kFb2PerLane * nsplit) {
      const int nhere = min(HW * kFb2PerLane, nfail - pass0);
      int ent[kFb2PerLane], c1[kFb2PerLane], c2[kFb2PerLane], c3[kFb2PerLane], el[kFb2PerLane];
      int eo[kFb2PerLane];  // the entry's index in the job's T arrays
      double I[kFb2PerLane], p[kFb2PerLane], res[kFb2PerLane];
      unsigned int done = 0;
#pragma unroll
      for (int j = 0; j < kFb2PerLane; j++) {
        const int k = sl + HW * j;
        ent[j] = -1;
        el[j] = 0;
        eo[j] = 0;
        int code = 0;
        if (gvalid && k < nhere) {
          int kk = pass0 + k, l = 0;
          while (kk >= gc[l]) kk -= gc[l++];
          const T2Group G = groups[g0 + l];
          ent[j] = flist[G.eoff + kk];
          code = tmpl[G.tmpl + ent[j]];
          el[j] = l;
          eo[j] = G.eoff + ent[j];
        }
        c1[j] = code & 0xff;
        c2[j] = (code >> 8) & 0xff;
        c3[j] = (code >> 16) & 0xff;
        if (ent[j] < 0) done |= 1u << j;
        I[j] = p[j] = res[j] = 0.0;
      }
      const unsigned int all = (1u << kFb2PerLane) - 1u;
      auto f = [&](int j, int slot) {
        double rn = 1.0;
        for (int k = 0; k < c3[j]; k++) rn = s_r[slot] * rn;
        return s_w[slot] *
               ((((s_cu[el[j] * 64 + slot] * rn) * s_ka[c1[j] * kKS + slot]) * s_kb[c2[j] * kKS + slot]) * s_ex[slot]);
      };
      __syncthreads();
      auto test = [&](int l, const double (&q)[kFb2PerLane]) {
        const double N = T.s92.lvl_n[l];
#pragma unroll
        for (int j = 0; j < kFb2PerLane; j++) {
          if (done & (1u << j)) continue;
          const double ee = I[j] - p[j];
          if (16 * ee * ee <= 3 * N * fabs(I[j] - q[j]) * T.tol) {
            done |= 1u << j;
            res[j] = 16 * I[j] / (3 * N);
          }
        }
      };
      // pair sums (0 + f(left)) + f(right) added to I in pair order; four
      // pairs' sums are formed before they are added, so that their LDS
      // reads overlap
      auto sum = [&](int l, int slot0, int kb, int ke) {
        const int b0 = T.s92.lvl_begin[l];
#pragma unroll
        for (int j = 0; j < kFb2PerLane; j++) {
          if (done & (1u << j)) continue;
          int k = kb;
          for (; k + 4 <= ke; k += 4) {
            double t[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              t[i] = f(j, slot0 + 2 * (k + i - kb));
              if (T.p92r[b0 + k + i] <= end) t[i] += f(j, slot0 + 2 * (k + i - kb) + 1);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) I[j] += t[i];
          }
          for (; k < ke; k++) {
            double t = f(j, slot0 + 2 * (k - kb));
            if (T.p92r[b0 + k] <= end) t += f(j, slot0 + 2 * (k - kb) + 1);
            I[j] += t;
          }
        }
      };
      int lv = 0;
      while (lv < T.s92.nlevels && wave_any(done != all)) {
        const int l1 = psm92_round_end(lv, T.s92.nlevels, HW);
        if (l1 > lv + 1 || psm92_points(lv) <= HW) {
          eval(base + sl, psm92_round_point(T, lv, l1, sl, end));
          __syncthreads();
          int roff = 0;
          for (int l = lv; l < l1; l++) {
            if (l == 0) {
#pragma unroll
              for (int j = 0; j < kFb2PerLane; j++)
                if (!(done & (1u << j))) I[j] = p[j] = f(j, base);
            } else {
              double q[kFb2PerLane];
#pragma unroll
              for (int j = 0; j < kFb2PerLane; j++) {
                q[j] = 2 * p[j];
                p[j] = 2 * I[j];
              }
              sum(l, base + roff, 0, T.s92.lvl_begin[l + 1] - T.s92.lvl_begin[l]);
              test(l, q);
            }
            roff += psm92_points(l);
          }
          __syncthreads();
          lv = l1;
        } else {
          double q[kFb2PerLane];
#pragma unroll
          for (int j = 0; j < kFb2PerLane; j++) {
            q[j] = 2 * p[j];
            p[j] = 2 * I[j];
          }
          const int b0 = T.s92.lvl_begin[lv], b1 = T.s92.lvl_begin[lv + 1];
          for (int cb = b0; cb < b1; cb += HW / 2) {
            const int nc = min(HW / 2, b1 - cb);
            const int k = sl >> 1;
            int n = -1;
            if (k < nc) {
              if ((sl & 1) == 0) n = T.p92l[cb + k];
              else if (T.p92r[cb + k] <= end) n = T.p92r[cb + k];
            }
            eval(base + sl, n);
            __syncthreads();
            sum(lv, base, cb - b0, cb - b0 + nc);
            __syncthreads();
          }
          test(lv, q);
          lv++;
        }
      }
#pragma unroll
      for (int j = 0; j < kFb2PerLane; j++) {
        if (ent[j] < 0) continue;
        if (done & (1u << j)) fbT[(size_t)eo[j] * J.maxPairs + mypair] = res[j];
        else if (fbf2) fbf2[eo[j]] = 1;
        else status[2 * t + 1] = 1;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused launches: one grid covers every class (lambda = la + lb for type 1,
// lambda_max = max(la, lb) + l for the type-2 fallback), each class owning a
// contiguous range of workgroups; the class bodies are those of the
// per-class kernels.  For small jobs, where each per-class launch holds only
// a few waves, this turns ~40 dependent launches into a handful.

struct ClassMap {
  int n;                     // classes with work
  int cls[16];               // class id of entry i
  int bbeg[17];              // workgroup range of entry i: [bbeg[i], bbeg[i+1])
};

__device__ __forceinline__ int class_entry(const ClassMap &cm, int b) {
  int i = 0;
  while (i + 1 < cm.n && b >= cm.bbeg[i + 1]) i++;
  return i;
}

// per-class views of the type-1 work arrays
struct T1Set {
  const T1Unit *units;
  double *q;
  unsigned long long *qf;
  int *fbq, *fbc, *tq;
  double *st;
  size_t plane;
  unsigned char *stat, *fbflag;
  int ubeg[ECP_MAX_LAB + 1], ucnt[ECP_MAX_LAB + 1];
  int head[ECP_MAX_LAB + 1];  // PS93 levels of the head per class (>= the level count: no tail)
  // speculative fallback (nullptr: off): open entries after the head, the
  // fallback's results and its failed-entry masks
  unsigned long long *qopen, *qfbf;
  double *qfb;
};

#define ECP_CASES_LAB(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#define ECP_CASES_LMB(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

template <int MAXLAB, int MODE>
__global__ __launch_bounds__(64) void k_type1_fused(DevTables T, DevJob J, T1Set S, ClassMap cm) {
  const int i = class_entry(cm, blockIdx.x);
  const int c = cm.cls[i], bid = blockIdx.x - cm.bbeg[i], nb = cm.bbeg[i + 1] - cm.bbeg[i];
  const int ub = S.ubeg[c];
  switch (c) {
#define X(L)                                                                                                   \
  case L:                                                                                                      \
    if constexpr (L <= MAXLAB)                                                                                 \
      type1_body<L, MODE>(bid, nb, T, J, S.units + ub, S.ucnt[c], S.q, S.qf + ub, S.fbq + ub, S.fbc + c, S.head[c], \
                          S.st, S.plane, S.tq + ub, S.fbc + 32 + c, S.fbflag, S.qopen ? S.qopen + ub : nullptr); \
    break;
    ECP_CASES_LAB(X)
#undef X
  }
}

template <int MAXLAB>
__global__ __launch_bounds__(64) void k_t1tail_fused(DevTables T, DevJob J, T1Set S, ClassMap cm) {
  __shared__ double s_T[nq_of(MAXLAB) * kTS];
  const int i = class_entry(cm, blockIdx.x);
  const int c = cm.cls[i], bid = blockIdx.x - cm.bbeg[i], nb = cm.bbeg[i + 1] - cm.bbeg[i];
  const int ub = S.ubeg[c];
  switch (c) {
#define X(L)                                                                                                   \
  case L:                                                                                                      \
    if constexpr (L <= MAXLAB)                                                                                 \
      t1tail_body<L>(bid, nb, T, J, S.units + ub, S.q, S.qf + ub, S.fbq + ub, S.fbc + c, S.head[c], S.st, S.plane, \
                     S.tq + ub, S.fbc + 32 + c, S.fbflag, s_T, S.qfb != nullptr);                                \
    break;
    ECP_CASES_LAB(X)
#undef X
  }
}

template <int MAXLAB, bool SPEC>
__global__ __launch_bounds__(64) void k_fb1_fused(DevTables T, DevJob J, T1Set S, ClassMap cm) {
  __shared__ double s_r[64], s_w[64], s_u[64], s_g[64], s_k[(MAXLAB + 1) * kKS], s_t[nq_of(MAXLAB) * kFb1TS];
  __shared__ int s_e[64];
  const int i = class_entry(cm, blockIdx.x);
  const int c = cm.cls[i], bid = blockIdx.x - cm.bbeg[i], nb = cm.bbeg[i + 1] - cm.bbeg[i];
  const int ub = S.ubeg[c];
  switch (c) {
#define X(L)                                                                                                   \
  case L:                                                                                                      \
    if constexpr (L <= MAXLAB)                                                                                 \
      fb1_body<L, SPEC>(bid, nb, T, J, S.units + ub, S.q, S.qf + ub, S.fbq + ub, S.fbc + c, S.stat, s_r, s_w, s_u, \
                        s_g, s_k, s_t, s_e, S.tq + ub, S.fbc + 32 + c, SPEC ? S.qopen + ub : nullptr, S.qfb,      \
                        SPEC ? S.qfbf + ub : nullptr);                                                          \
    break;
    ECP_CASES_LAB(X)
#undef X
  }
}

struct T2Set {
  const T2Group *groups;
  const int *tmpl;
  const unsigned char *tf;
  const int *fbq2;  // per class, the triplets with failed entries (nt apart)
  const int *fbc;
  double *fbT;
  unsigned char *stat;
  int nt;
  const int *gcount, *flist, *tgfirst;
  int nsplit;
  unsigned char *fbf2;  // speculative fallback: per entry, the large grid failed (nullptr: off)
};

// PACK: triplets with few failed entries run several pairs per wave (three
// waves per SIMD kept: at two the uneven batch loses the gain)
template <int MAXLMB, bool PACK>
__global__ __launch_bounds__(64, PACK ? 3 : 1) void k_fb2_fused(DevTables T, DevJob J, T2Set S, ClassMap cm) {
  __shared__ double s_r[64], s_w[64], s_cu[ECP_MAX_L * 64], s_ex[64];
  __shared__ double s_ka[(MAXLMB + 1) * kKS], s_kb[(MAXLMB + 1) * kKS];
  const int i = class_entry(cm, blockIdx.x);
  const int c = cm.cls[i], bid = blockIdx.x - cm.bbeg[i], nb = cm.bbeg[i + 1] - cm.bbeg[i];
  switch (c) {
#define X(L)                                                                                                   \
  case L:                                                                                                      \
    if constexpr (L <= MAXLMB)                                                                                 \
      fb2_body<L, PACK>(bid, nb, T, J, S.groups, S.tmpl, S.tf, S.fbq2 + (size_t)L * S.nt, S.fbc + kFb2Cnt + L, S.fbT, \
                  S.stat, S.gcount, S.flist, S.tgfirst, S.nsplit, s_r, s_w, s_cu, s_ex, s_ka, s_kb, S.fbf2);    \
    break;
    ECP_CASES_LMB(X)
#undef X
  }
}

// ---------------------------------------------------------------------------
// assembly: type-1 block (calcChi), type-2 block (calcGamma), binomial shift
// (calcPolynomials) and norms (libecp.c:234-235)

constexpr int kAsmThreads = 64;
constexpr int kAsmPack = 4, kAsmPackLab = 2;  // packed type-1 assembly: triplets per wave, largest la + lb

struct AsmArgs {
  const T1Unit *units1;
  const int *t1first;        // per triplet: first type-1 unit
  const double *qout;
  const T2Group *groups;
  const int *tgfirst;        // per triplet: first group
  const int *tmpl;
  const double *Tout;
  const unsigned char *tfail;
  const unsigned char *fbf2;  // speculative type-2 fallback: per entry, the large grid failed (nullptr: off)
  const double *fbT;
  const double *fac, *dfac;
  double *blocks;
  // dynamic LDS layout (doubles): J is shared; the type-1 region (rsph, H)
  // and the type-2 region (blk, T) overlap
  int lds_blk, lds_H, lds_J, lds_rsph, lds_T, lds_rs_stride, pair_chunk, lds_Q;
  int lds_S1, lds_S2, shift_lm;  // shift factor tables (aliasing rsph/Q resp. T) for shells up to shift_lm
  const double *udir;             // type 1: per unit, cos theta, phi, sin phi, cos phi of its pair direction
  int lds_group;                  // packed type-1 assembly: LDS doubles per triplet
  // speculative type-1 fallback (qfb != nullptr): an entry whose small grid
  // failed (qfail bit) takes the fallback's result qfb; the triplet failed
  // where the fallback failed too (qfbf bit)
  const unsigned long long *qfail, *qfbf;
  const double *qfb;
};

// binomial shift of a block over C-centred monomials to the IJK(la) x IJK(lb)
// output (util.c:246-334).  The block is read through get(r, q): row r is an
// ordering position of degree <= la, column q a position of degree <= lb.
struct V3 {
  double x, y, z;
};

// Nonzero binomial factors of the shift (util.c:246-334), listed once per
// triplet: for component c of shell l (ordering position), the factors
// C(a,alpha) (C-A)^(a-alpha) over alpha in the loop order of util.c:277-320,
// each formed exactly as the per-element loops would (powers by repeated
// multiplication, binomials left to right, factors |f| <= accuracy dropped,
// the norm applied to the B side), with the position of alpha.
__host__ __device__ constexpr int shift_mt(int l) {
  int best = 1;
  for (int ax = 0; ax <= l; ax++)
    for (int ay = 0; ax + ay <= l; ay++) {
      const int v = (ax + 1) * (ay + 1) * (l - ax - ay + 1);
      best = v > best ? v : best;
    }
  return best;
}

// doubles of LDS for the factor tables of both sides for shells up to LM
__host__ __device__ constexpr int shift_tab_doubles(int LM) {
  const int n = ECP_DIM_IJK(LM) * shift_mt(LM);
  return 2 * (n + (n + 1) / 2 + (ECP_DIM_IJK(LM) + 1) / 2);
}

struct ShiftTab {
  double *f;
  int *r, *n;
  int mt;
};

__device__ __forceinline__ ShiftTab shift_tab_at(double *base, int LM, int side) {
  const int n = ECP_DIM_IJK(LM) * shift_mt(LM);
  double *b = base + side * (n + (n + 1) / 2 + (ECP_DIM_IJK(LM) + 1) / 2);
  ShiftTab S;
  S.f = b;
  S.r = (int *)(b + n);
  S.n = (int *)(b + n + (n + 1) / 2);
  S.mt = shift_mt(LM);
  return S;
}

__device__ void shift_factors(const DevTables &T, int l, const V3 v, bool with_norm, double norm, ShiftTab S,
                              int lid, int nlanes) {
  const int D = T.ijkDim;
  const double tol = T.accuracy;
  for (int c = lid; c < ECP_DIM_IJK(l); c += nlanes) {
    const int *e = T.ijk + 3 * ECP_IDX_C(l, c);
    const int ax = e[0], ay = e[1], az = e[2];
    int n = 0;
    for (int alx = 0; alx <= ax; alx++) {
      double x = 1.0;
      for (int i = 0; i < ax - alx; i++) x = -x * v.x;
      const double bx = T.binom[ax * 16 + alx];
      for (int aly = 0; aly <= ay; aly++) {
        double y = 1.0;
        for (int i = 0; i < ay - aly; i++) y = -y * v.y;
        const double by = bx * T.binom[ay * 16 + aly];
        for (int alz = 0; alz <= az; alz++) {
          double z = 1.0;
          for (int i = 0; i < az - alz; i++) z = -z * v.z;
          const double bz = by * T.binom[az * 16 + alz];
          double factor = bz * ((x * y) * z);
          if (fabs(factor) <= tol) continue;
          if (with_norm) factor *= norm;
          S.f[c * S.mt + n] = factor;
          S.r[c * S.mt + n] = T.ijkIndex[(alx * D + aly) * D + alz];
          n++;
        }
      }
    }
    S.n[c] = n;
  }
}

// binomial shift of a block over C-centred monomials to the IJK(la) x IJK(lb)
// output (util.c:246-334).  The block is read through get(r, q): row r is an
// ordering position of degree <= la, column q a position of degree <= lb.
// tab: LDS for the factor tables (shift_tab_doubles(LM) doubles, LM >= la,
// lb), free to overwrite; Jbuf: LDS for the half-shifted block.  The lanes
// lid = 0 .. nlanes - 1 of the workgroup do the work; every lane of the
// workgroup calls it (it synchronises the workgroup), those of an inactive
// lane group with active = false.
template <typename Get>
__device__ void shift_block(const DevTables &T, Get get, int la, int lb, const V3 rA, const V3 rB, double norm,
                            double *Jbuf, double *tab, int LM, double *out, int lid = -1, int nlanes = 0,
                            bool active = true) {
  if (lid < 0) {
    lid = threadIdx.x;
    nlanes = blockDim.x;
  }
  const int ncb = ECP_DIM_C(lb), nia = active ? ECP_DIM_IJK(la) : 0, nib = ECP_DIM_IJK(lb);
  const ShiftTab SA = shift_tab_at(tab, LM, 0), SB = shift_tab_at(tab, LM, 1);
  __syncthreads();
  if (active) {
    shift_factors(T, la, rA, false, 1.0, SA, lid, nlanes);
    shift_factors(T, lb, rB, true, norm, SB, lid, nlanes);
  }
  __syncthreads();
  for (int idx = lid; idx < nia * ncb; idx += nlanes) {
    const int c1 = idx / ncb, q = idx % ncb;
    const double *f = SA.f + c1 * SA.mt;
    const int *r = SA.r + c1 * SA.mt;
    const int n = SA.n[c1];
    double acc = 0.0;
    for (int t = 0; t < n; t++) acc += f[t] * get(r[t], q);
    Jbuf[idx] = acc;
  }
  __syncthreads();
  for (int idx = lid; idx < nia * nib; idx += nlanes) {
    const int c1 = idx / nib, c2 = idx % nib;
    const double *f = SB.f + c2 * SB.mt;
    const int *r = SB.r + c2 * SB.mt;
    const int n = SB.n[c2];
    double acc = 0.0;
    for (int t = 0; t < n; t++) acc += f[t] * Jbuf[c1 * ncb + r[t]];
    out[idx] = acc;
  }
  __syncthreads();
}

// H[p] += sum over the chunk's pairs k, in pair order, of
// sum_{l = D, D - 2, ..} (sum_m S_lm(k) poly2sph[p][lm]) Q_k[(D - l) / 2]
// for a monomial p of degree D: the same terms in the same order as the
// general loop in k_asm1, with the monomial's poly2sph row held in
// registers across the pairs instead of read again for every pair.
template <int D>
__device__ __forceinline__ double asm1_monomial(double acc, const double *__restrict__ pm, const double *rsph,
                                                int rs_stride, const double *Qs, int nql, int nc) {
  constexpr int NPM = (D + 1) * (D + 2) / 2;  // sum of 2 l + 1 over l = D, D - 2, ..
  double pmv[NPM];
  {
    int j = 0;
#pragma unroll
    for (int l = D; l >= 0; l -= 2)
#pragma unroll
      for (int m = 0; m < 2 * l + 1; m++) pmv[j++] = pm[ECP_IDX_LM(l, m)];
  }
  for (int k = 0; k < nc; k++) {
    const double *rs = rsph + k * rs_stride;
    const double *Q = Qs + k * nql;
    int j = 0;
#pragma unroll
    for (int l = D; l >= 0; l -= 2) {
      double f = 0.0;
#pragma unroll
      for (int m = 0; m < 2 * l + 1; m++) f += rs[ECP_IDX_LM(l, m)] * pmv[j++];
      acc += f * Q[(D - l) / 2];
    }
  }
  return acc;
}

// Type-1 block of one triplet per wave (the triplets tlist[0 .. ntl), in
// class order lambda = la + lb).  chi[i][j] depends on (i, j) only
// through the monomial p = i + j, so H[p] = sum_pairs sum_l F_pair[p][l]
// Q[deg p][l] with F_pair[p][l] = sum_m S_lm(k^_pair) poly2sph[p][lm] is
// formed once per monomial (same terms, same order as type1.c:266-295 per
// element), then shifted to the output block (util.c:246-334).
// G > 1 (classes la + lb <= 2, where a triplet has at most 10 monomials):
// G triplets per wave, kAsmThreads / G lanes and an LDS slice of
// A.lds_group doubles each; the wave's lane groups run the same phases and
// synchronise together, an inactive group with empty loops.
// MAXD: the largest monomial degree whose poly2sph row goes to registers
// (the triplets' la + lb when a launch covers one class); higher degrees
// take the general loop.  Registers, without spills: 51 (rows up to degree
// 2 in registers) to 111 (degree 6), 94 for the packed variant; the pair
// directions come from the tables kernel, so no trigonometric code sits
// here.
template <int G = 1, int MAXD = 6>
__global__ __launch_bounds__(kAsmThreads, 3) void k_asm1(DevTables T, DevJob J, AsmArgs A,
                                                      unsigned char *__restrict__ status,
                                                      const int *__restrict__ tlist, int ntl,
                                                      const unsigned char *__restrict__ fbflag, int filter) {
  extern __shared__ double lds_all[];
  constexpr int W = kAsmThreads / G;
  const int grp = G == 1 ? 0 : (int)threadIdx.x / W;
  const int lid = G == 1 ? (int)threadIdx.x : (int)threadIdx.x % W;
  double *lds = lds_all + (size_t)grp * A.lds_group;
  double *H = lds + A.lds_H;
  double *Jbuf = lds + A.lds_J;
  double *rsph = lds + A.lds_rsph;
  double *Qs = lds + A.lds_Q;
  int t;
  bool active = true;
  if constexpr (G == 1) {
    t = tlist[blockIdx.x];
    // filter 1: only triplets without large-grid work, 2: only those with it
    if ((filter == 1 && fbflag[t]) || (filter == 2 && !fbflag[t])) return;
    if (status[2 * t]) return;
  } else {
    const int ti = blockIdx.x * G + grp;
    active = ti < ntl;
    t = active ? tlist[ti] : tlist[0];
    if (active && ((filter == 1 && fbflag[t]) || (filter == 2 && !fbflag[t]) || status[2 * t])) active = false;
    if (!wave_any(active)) return;
  }
  const ecp_triplet tr = J.trip[t];
  const ecp_centre Cn = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const int la = tr.la, lb = tr.lb, lab = la + lb;  // shifted for derivatives
  if (A.qfb) {
    // speculative fallback: an integral failed where the small grid and
    // the large grid both failed (type1.c:182-190, code 1)
    bool bad = false;
    if (active)
      for (int k = lid; k < Sa.K * Sb.K; k += W) {
        const int ug = A.t1first[t] + k;
        bad = bad || (A.qfail[ug] & A.qfbf[ug]) != 0ull;
      }
    if constexpr (G == 1) {
      if (wave_any(bad)) {
        if (lid == 0) status[2 * t] = 1;
        return;
      }
    } else {
      // the decision of the lane group's W lanes
      const unsigned long long m = __ballot(bad);
      if ((m >> (grp * W)) & ((1ull << W) - 1ull)) {
        if (lid == 0) status[2 * t] = 1;
        active = false;
      }
      if (!wave_any(active)) return;
    }
  }
  const ecp_xyz rc = J.atoms[Cn.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const V3 vA{ra.x - rc.x, ra.y - rc.y, ra.z - rc.z}, vB{rb.x - rc.x, rb.y - rc.y, rb.z - rc.z};
  const int D = T.ijkDim;
  const int tmd = ECP_DIM_L(T.tmDim);

  const int np = active ? ECP_DIM_C(lab) : 0;
  const int nql = nq_of(lab);
  const int npairs = active ? Sa.K * Sb.K : 0;
  int npmax = npairs;
  if (G > 1)
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
  const int u0 = A.t1first[t];
  int qstart = 0;
  for (int idx = lid; idx < np; idx += W) H[idx] = 0.0;
  for (int c0 = 0; c0 < npmax; c0 += A.pair_chunk) {
    const int nc = G == 1 ? min(A.pair_chunk, npairs - c0) : max(0, min(A.pair_chunk, npairs - c0));
    // the harmonics of the chunk's pair directions (formed by the tables
    // kernel), one (pair, m) column per lane; the pairs' Q values staged
    // alongside
    for (int it = lid; it < nc * (lab + 1); it += W) {
      const int k = it / (lab + 1), m = it - k * (lab + 1);
      const double *d = A.udir + 4 * (size_t)(u0 + c0 + k);
      real_sph_column(lab, m, d[0], d[1], d[2], d[3], A.fac, A.dfac, rsph + k * A.lds_rs_stride);
    }
    for (int it = lid; it < nc * nql; it += W) {
      const int k = it / nql, e = it - k * nql;
      const int ug = u0 + c0 + k;
      const int qo = A.units1[ug].qoff + e;
      Qs[k * nql + e] = (A.qfb && ((A.qfail[ug] >> e) & 1ull)) ? A.qfb[qo] : A.qout[qo];
    }
    __syncthreads();
    for (int p = lid; p < np; p += W) {
      const int *e = T.ijk + 3 * p;
      const int deg = e[0] + e[1] + e[2];
      qstart = 0;
      for (int l1 = 0; l1 < deg; l1++) qstart += l1 / 2 + 1;
      const double *pm = T.poly2sph + (size_t)p * tmd;
      double acc = H[p];
      if (deg <= MAXD) {
        switch (deg) {
#define MONO(D)                                                                       \
  case D:                                                                             \
    if constexpr (D <= MAXD) acc = asm1_monomial<D>(acc, pm, rsph, A.lds_rs_stride, Qs + qstart, nql, nc); \
    break;
          MONO(0) MONO(1) MONO(2) MONO(3) MONO(4) MONO(5) MONO(6)
#undef MONO
        }
        H[p] = acc;
        continue;
      }
      for (int k = 0; k < nc; k++) {
        const double *rs = rsph + k * A.lds_rs_stride;
        const double *Q = Qs + k * nql + qstart;
        for (int l = deg; l >= 0; l -= 2) {
          double f = 0.0;
          f = dot_in_order<2>(f, rs + ECP_IDX_LM(l, 0), 1, pm + ECP_IDX_LM(l, 0), 1, 2 * l + 1);
          acc += f * Q[(deg - l) / 2];
        }
      }
      H[p] = acc;
    }
    __syncthreads();
  }
  const double norm1 = 4.0 * M_PI;
  auto get1 = [&](int r, int q) {
    const int *ea = T.ijk + 3 * r, *eb = T.ijk + 3 * q;
    return H[T.ijkIndex[((ea[0] + eb[0]) * D + ea[1] + eb[1]) * D + ea[2] + eb[2]]];
  };
  if constexpr (G == 1)
    shift_block(T, get1, tr.oa, tr.ob, vA, vB, norm1, Jbuf, lds + A.lds_S1, A.shift_lm, A.blocks + tr.blk1);
  else
    shift_block(T, get1, tr.oa, tr.ob, vA, vB, norm1, Jbuf, lds + A.lds_S1, A.shift_lm, A.blocks + tr.blk1, lid, W,
                active);
}

// Type-2 block of one triplet per wave: for each projector l, the T_l
// entries (small-grid results, or the sum of the per-pair fallback results)
// are placed in LDS and contracted with the angular tables of both atoms
// (type2.c:583-623), then the block is shifted (util.c:246-334).
// NB > 0: the angular factors of atom B for every lambda2 (at most NB
// values) and m (at most NM) are held in registers for each l, those of atom
// A per lambda1; NB = 0: both read for each (lambda1, lambda2), with NM > 1
// the 2 l + 1 values of each loaded together (NM >= 2 l + 1 of every
// projector l).
// tlist: the triplets of the launch, one per workgroup (nullptr: workgroup
// t takes triplet t).
template <int NB, int NM, int TPB = kAsmThreads>
__global__ __launch_bounds__(TPB) void k_asm2(DevTables T, DevJob J, AsmArgs A,
                                                      unsigned char *__restrict__ status,
                                                      const unsigned char *__restrict__ fbflag, int filter,
                                                      const int *__restrict__ tlist) {
  extern __shared__ double lds[];
  double *blk = lds + A.lds_blk;
  double *Jbuf = lds + A.lds_J;
  double *Tl = lds + A.lds_T;
  const int t = tlist ? tlist[blockIdx.x] : (int)blockIdx.x;
  if ((filter == 1 && fbflag[t]) || (filter == 2 && !fbflag[t])) return;
  if (status[2 * t + 1]) return;
  const ecp_triplet tr = J.trip[t];
  const ecp_centre Cn = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const int la = tr.la, lb = tr.lb, lab = la + lb;  // shifted for derivatives
  const int nca = ECP_DIM_C(la), ncb = ECP_DIM_C(lb);
  const ecp_xyz rc = J.atoms[Cn.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const double rAC[3] = {ra.x - rc.x, ra.y - rc.y, ra.z - rc.z};
  const double rBC[3] = {rb.x - rc.x, rb.y - rc.y, rb.z - rc.z};
  const V3 vA{rAC[0], rAC[1], rAC[2]}, vB{rBC[0], rBC[1], rBC[2]};
  const int D = T.ijkDim;
  const int g0 = A.tgfirst[t];
  if (A.fbf2 && filter == 2) {
    // speculative fallback: the triplet failed where an entry failed on the
    // small grid and on the large grid (type2.c:417-528, code 1)
    bool bad = false;
    for (int l = 0; l < Cn.L; l++) {
      const T2Group G = A.groups[g0 + l];
      for (int e = threadIdx.x; e < G.ne; e += blockDim.x) bad = bad || (A.tfail[G.eoff + e] && A.fbf2[G.eoff + e]);
    }
    if (__syncthreads_or(bad)) {
      if (threadIdx.x == 0) status[2 * t + 1] = 1;
      return;
    }
  }

  for (int i = threadIdx.x; i < nca * ncb; i += blockDim.x) blk[i] = 0.0;
  __syncthreads();
  const int nlm2 = J.angL * J.angL;
  const double *angA = J.ang + Cn.ang_off + (size_t)(Sa.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
  const double *angB = J.ang + Cn.ang_off + (size_t)(Sb.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
  for (int l = 0; l < Cn.L; l++) {
    const T2Group G = A.groups[g0 + l];
    const int d1 = la + l + 1, d2 = lb + l + 1, d3 = lab + 1;
    for (int i = threadIdx.x; i < d1 * d2 * d3; i += blockDim.x) Tl[i] = 0.0;
    __syncthreads();
    const int nm = 2 * l + 1;
    for (int e = threadIdx.x; e < G.ne; e += blockDim.x) {
      const int code = A.tmpl[G.tmpl + e];
      const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
      double v;
      if (!A.tfail[G.eoff + e]) {
        v = A.Tout[G.eoff + e];
      } else {
        v = sum_in_order(0.0, A.fbT + (size_t)(G.eoff + e) * J.maxPairs, Sa.K * Sb.K);
      }
      Tl[(l1 * d2 + l2) * d3 + l3] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nca * ncb; idx += blockDim.x) {
      const int pp = idx / ncb, qq = idx % ncb;
      int alpha = 0, beta = 0;
      while (ECP_DIM_C(alpha) <= pp) alpha++;
      while (ECP_DIM_C(beta) <= qq) beta++;
      const int par1 = (alpha + l) % 2, par2 = (beta + l) % 2;
      const int ll1 = (par1 > l - alpha) ? par1 : l - alpha;
      const int ll2 = (par2 > l - beta) ? par2 : l - beta;
      double tmp = 0.0;
      if constexpr (NB > 0) {
        double Bv[NB][NM];
#pragma unroll
        for (int i = 0; i < NB; i++) {
          const int lam2 = ll2 + 2 * i;
          const double *pb = angB + ((size_t)lam2 * nlm2 + l * l) * J.ang_c + qq;
#pragma unroll
          for (int m = 0; m < NM; m++) Bv[i][m] = (lam2 <= lb + l && m < nm) ? pb[m * J.ang_c] : 0.0;
        }
        for (int lam1 = ll1; lam1 <= la + l; lam1 += 2) {
          const double *pa = angA + ((size_t)lam1 * nlm2 + l * l) * J.ang_c + pp;
          double Av[NM];
#pragma unroll
          for (int m = 0; m < NM; m++) Av[m] = m < nm ? pa[m * J.ang_c] : 0.0;
#pragma unroll
          for (int i = 0; i < NB; i++) {
            const int lam2 = ll2 + 2 * i;
            if (lam2 <= lb + l) {
              double factor = 0.0;
#pragma unroll
              for (int m = 0; m < NM; m++)
                if (m < nm) factor += Av[m] * Bv[i][m];
              tmp += factor * Tl[(lam1 * d2 + lam2) * d3 + alpha + beta];
            }
          }
        }
      } else if constexpr (NM > 1) {
        // all 2 l + 1 factor pairs of an angular factor loaded together (NM
        // >= 2 l + 1), then summed in m order
        for (int lam1 = ll1; lam1 <= la + l; lam1 += 2)
          for (int lam2 = ll2; lam2 <= lb + l; lam2 += 2) {
            const double *pa = angA + ((size_t)lam1 * nlm2 + l * l) * J.ang_c + pp;
            const double *pb = angB + ((size_t)lam2 * nlm2 + l * l) * J.ang_c + qq;
            double x[NM], y[NM];
#pragma unroll
            for (int m = 0; m < NM; m++) {
              x[m] = m < nm ? pa[m * J.ang_c] : 0.0;
              y[m] = m < nm ? pb[m * J.ang_c] : 0.0;
            }
            double factor = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++)
              if (m < nm) factor += x[m] * y[m];
            tmp += factor * Tl[(lam1 * d2 + lam2) * d3 + alpha + beta];
          }
      } else {
        for (int lam1 = ll1; lam1 <= la + l; lam1 += 2)
          for (int lam2 = ll2; lam2 <= lb + l; lam2 += 2) {
            const double *pa = angA + ((size_t)lam1 * nlm2 + l * l) * J.ang_c + pp;
            const double *pb = angB + ((size_t)lam2 * nlm2 + l * l) * J.ang_c + qq;
            const double factor = dot_in_order<2>(0.0, pa, J.ang_c, pb, J.ang_c, nm);
            tmp += factor * Tl[(lam1 * d2 + lam2) * d3 + alpha + beta];
          }
      }
      blk[idx] += tmp;
    }
    __syncthreads();
  }
  const double norm2 = (4.0 * M_PI) * (4.0 * M_PI);
  auto get2 = [&](int r, int q) { return blk[r * ncb + q]; };
  shift_block(T, get2, tr.oa, tr.ob, vA, vB, norm2, Jbuf, lds + A.lds_S2, A.shift_lm, A.blocks + tr.blk2);
}

// Type-2 blocks of classes la + lb <= kAsmPackLab (at most 9 elements),
// kAsmPack triplets per wave (tlist[0 .. ntl)): lane groups of
// kAsmThreads / kAsmPack lanes, each with an LDS slice of A.lds_group
// doubles, run k_asm2's phases side by side with the same arithmetic (NM as
// there: the 2 l + 1 angular factors loaded together, or dot_in_order for
// NM = 1).  The channel loop runs to the wave's largest L, a group past its
// own with empty loops; every group synchronises with the wave.
template <int NM>
__global__ __launch_bounds__(kAsmThreads) void k_asm2p(DevTables T, DevJob J, AsmArgs A,
                                                    unsigned char *__restrict__ status,
                                                    const int *__restrict__ tlist, int ntl,
                                                    const unsigned char *__restrict__ fbflag, int filter) {
  extern __shared__ double lds_all[];
  constexpr int G = kAsmPack, W = kAsmThreads / G;
  const int grp = (int)threadIdx.x / W, lid = (int)threadIdx.x % W;
  double *lds = lds_all + (size_t)grp * A.lds_group;
  double *blk = lds + A.lds_blk;
  double *Jbuf = lds + A.lds_J;
  double *Tl = lds + A.lds_T;
  const int ti = blockIdx.x * G + grp;
  bool active = ti < ntl;
  const int t = active ? tlist[ti] : tlist[0];
  if (active && ((filter == 1 && fbflag[t]) || (filter == 2 && !fbflag[t]) || status[2 * t + 1])) active = false;
  if (!wave_any(active)) return;
  const ecp_triplet tr = J.trip[t];
  const ecp_centre Cn = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const int la = tr.la, lb = tr.lb, lab = la + lb;  // shifted for derivatives
  const int nca = ECP_DIM_C(la), ncb = ECP_DIM_C(lb);
  const ecp_xyz rc = J.atoms[Cn.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const V3 vA{ra.x - rc.x, ra.y - rc.y, ra.z - rc.z}, vB{rb.x - rc.x, rb.y - rc.y, rb.z - rc.z};
  const int g0 = A.tgfirst[t];
  if (A.fbf2 && filter == 2) {
    // speculative fallback: the triplet failed where an entry failed on the
    // small grid and on the large grid (type2.c:417-528, code 1); the
    // decision of the group's lanes
    bool bad = false;
    if (active)
      for (int l = 0; l < Cn.L; l++) {
        const T2Group Gr = A.groups[g0 + l];
        for (int e = lid; e < Gr.ne; e += W) bad = bad || (A.tfail[Gr.eoff + e] && A.fbf2[Gr.eoff + e]);
      }
    const unsigned long long m = __ballot(bad);
    if ((m >> (grp * W)) & ((1ull << W) - 1ull)) {
      if (lid == 0) status[2 * t + 1] = 1;
      active = false;
    }
    if (!wave_any(active)) return;
  }
  if (active)
    for (int i = lid; i < nca * ncb; i += W) blk[i] = 0.0;
  int Lw = active ? Cn.L : 0;
  for (int o = 32; o > 0; o >>= 1) Lw = max(Lw, __shfl_xor(Lw, o, 64));
  __syncthreads();
  const int nlm2 = J.angL * J.angL;
  const double *angA = J.ang + Cn.ang_off + (size_t)(Sa.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
  const double *angB = J.ang + Cn.ang_off + (size_t)(Sb.atom - Cn.atom0) * J.ang_lam * nlm2 * J.ang_c;
  for (int l = 0; l < Lw; l++) {
    const bool on = active && l < Cn.L;
    const int d1 = la + l + 1, d2 = lb + l + 1, d3 = lab + 1;
    T2Group Gr{};
    if (on) {
      Gr = A.groups[g0 + l];
      for (int i = lid; i < d1 * d2 * d3; i += W) Tl[i] = 0.0;
    }
    __syncthreads();
    const int nm = 2 * l + 1;
    if (on)
      for (int e = lid; e < Gr.ne; e += W) {
        const int code = A.tmpl[Gr.tmpl + e];
        const int l1 = code & 0xff, l2 = (code >> 8) & 0xff, l3 = (code >> 16) & 0xff;
        double v;
        if (!A.tfail[Gr.eoff + e]) {
          v = A.Tout[Gr.eoff + e];
        } else {
          v = sum_in_order(0.0, A.fbT + (size_t)(Gr.eoff + e) * J.maxPairs, Sa.K * Sb.K);
        }
        Tl[(l1 * d2 + l2) * d3 + l3] = v;
      }
    __syncthreads();
    if (on)
      for (int idx = lid; idx < nca * ncb; idx += W) {
        const int pp = idx / ncb, qq = idx % ncb;
        int alpha = 0, beta = 0;
        while (ECP_DIM_C(alpha) <= pp) alpha++;
        while (ECP_DIM_C(beta) <= qq) beta++;
        const int par1 = (alpha + l) % 2, par2 = (beta + l) % 2;
        const int ll1 = (par1 > l - alpha) ? par1 : l - alpha;
        const int ll2 = (par2 > l - beta) ? par2 : l - beta;
        double tmp = 0.0;
        for (int lam1 = ll1; lam1 <= la + l; lam1 += 2)
          for (int lam2 = ll2; lam2 <= lb + l; lam2 += 2) {
            const double *pa = angA + ((size_t)lam1 * nlm2 + l * l) * J.ang_c + pp;
            const double *pb = angB + ((size_t)lam2 * nlm2 + l * l) * J.ang_c + qq;
            double factor;
            if constexpr (NM > 1) {
              double x[NM], y[NM];
#pragma unroll
              for (int m = 0; m < NM; m++) {
                x[m] = m < nm ? pa[m * J.ang_c] : 0.0;
                y[m] = m < nm ? pb[m * J.ang_c] : 0.0;
              }
              factor = 0.0;
#pragma unroll
              for (int m = 0; m < NM; m++)
                if (m < nm) factor += x[m] * y[m];
            } else {
              factor = dot_in_order<2>(0.0, pa, J.ang_c, pb, J.ang_c, nm);
            }
            tmp += factor * Tl[(lam1 * d2 + lam2) * d3 + alpha + beta];
          }
        blk[idx] += tmp;
      }
    __syncthreads();
  }
  const double norm2 = (4.0 * M_PI) * (4.0 * M_PI);
  auto get2 = [&](int r, int q) { return blk[r * ncb + q]; };
  shift_block(T, get2, tr.oa, tr.ob, vA, vB, norm2, Jbuf, lds + A.lds_S2, A.shift_lm, A.blocks + tr.blk2, lid, W,
              active);
}

// ---------------------------------------------------------------------------
// host-side orchestration

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// The per-job tables in one launch of 128-thread blocks: the zero-fill of
// per-run counters and flags (a kernel, so that the sequence can be captured
// in a HIP graph without memset nodes), the U tables per (centre, row, point
// chunk), the F slabs per (centre, shell, run) and the pair directions of
// the type-1 units.  They are independent of each other; every consumer is
// a later launch.
struct TabArgs {
  unsigned int *zero_base;
  size_t zero_words;
  int nzb, nub, nfit, rows;  // rows: U-table rows per centre (largest L + 1)
  const int2 *fit;
  const T1Unit *units;       // type-1 units and their pair directions, 128 per workgroup
  int nunits;
  double *udir;
};

// The pair direction k = 2 (za rAC + zb rBC) of a type-1 unit (type1.c:
// 211-302 via util.c:74-106), stored as cos theta, phi, sin phi, cos phi,
// the values spherical_harmonics.c:15-114 forms from it.  Formed with the
// tables, so that the type-1 assembly carries no trigonometric code (with
// it, 168 registers and spills instead of 111).
__device__ void unit_dir(const DevJob &J, const T1Unit &U, double *out) {
  const ecp_triplet tr = J.trip[U.trip];
  const ecp_centre Cn = J.cent[tr.centre];
  const ecp_shell Sa = J.shells[tr.sa], Sb = J.shells[tr.sb];
  const ecp_xyz rc = J.atoms[Cn.atom], ra = J.atoms[Sa.atom], rb = J.atoms[Sb.atom];
  const double za = J.pa[Sa.prim + U.pa], zb = J.pa[Sb.prim + U.pb];
  double r, th, ph;
  sph_coords(2.0 * (za * (ra.x - rc.x) + zb * (rb.x - rc.x)), 2.0 * (za * (ra.y - rc.y) + zb * (rb.y - rc.y)),
             2.0 * (za * (ra.z - rc.z) + zb * (rb.z - rc.z)), r, th, ph);
  out[0] = cos(th);
  out[1] = ph;
  out[2] = sin(ph);
  out[3] = cos(ph);
}

constexpr int kUtabChunks = (SO + 127) / 128;

__global__ __launch_bounds__(128, 4) void k_tables(DevTables T, DevJob J, TabArgs A) {
  int b = blockIdx.x;
  if (b < A.nzb) {
    for (size_t i = (size_t)b * 128 + threadIdx.x; i < A.zero_words; i += (size_t)A.nzb * 128) A.zero_base[i] = 0u;
    return;
  }
  b -= A.nzb;
  if (b < A.nub) {
    const int per_c = A.rows * kUtabChunks;
    const int c = b / per_c, rem = b - c * per_c;
    utab_rows(T, J, c, rem / kUtabChunks, rem % kUtabChunks);
    return;
  }
  b -= A.nub;
  if (b < A.nfit) {
    ftab_item(T, J, A.fit[b]);
    return;
  }
  b -= A.nfit;
  const int u = b * 128 + threadIdx.x;
  if (u < A.nunits) unit_dir(J, A.units[u], A.udir + 4 * (size_t)u);
}

// angular tables: a launch of their own (few registers, high occupancy),
// needed only by the type-2 assembly
__global__ __launch_bounds__(128) void k_ang(DevTables T, DevJob J, const int2 *__restrict__ items,
                                             const double *__restrict__ fac, const double *__restrict__ dfac) {
  __shared__ double rsph[kAngRsph];
  ang_item(T, J, items[blockIdx.x], fac, dfac, rsph);
}

// byte offsets of a job's arrays inside its device block: inputs (uploaded)
// first, then device-only buffers
struct JobLayout {
  size_t i_atoms = 0, i_amaxl = 0, i_sh = 0, i_pa = 0, i_pd = 0, i_cent = 0, i_gauss = 0, i_win = 0, i_trip = 0,
         i_u1 = 0, i_t1f = 0, i_grp = 0, i_goff = 0, i_tgf = 0, i_fit = 0, i_ait = 0, i_tcls = 0, i_tcls2 = 0,
         in_bytes = 0;
  size_t d_zero = 0, d_fbc = 0, d_gflag = 0, d_stat = 0, d_t1fb = 0, d_t2fb = 0, d_tclaim = 0, d_fbf2 = 0, d_zero_end = 0, d_UL = 0, d_Ut = 0, d_F = 0, d_ang = 0,
         d_q = 0, d_t1s = 0, d_tq1 = 0, d_qf = 0, d_qfb = 0, d_qopen = 0, d_qfbf = 0, d_udir = 0, d_fbq1 = 0, d_T = 0, d_tf = 0, d_flist = 0, d_tq2 = 0, d_fbq2 = 0, d_fbT = 0, d_blk = 0, total = 0;
  int FL = 0, UK = 0, angL = 0, ang_lam = 0, ang_c = 0, deriv = 0;
};

struct Prepared {
  int nt = 0;
  int ncent = 0;
  int nfit = 0, nait = 0;
  int cls_begin[ECP_MAX_LAB + 2] = {0};
  int cls_count[ECP_MAX_LAB + 1] = {0};
  int lmb_count[kMaxLMB + 1] = {0};
  int ngroups = 0, ntot2 = 0, maxPairs = 0;
  size_t nblk = 0;
  // device pointers
  DevJob J{};
  const T1Unit *u1 = nullptr;
  double *q = nullptr;
  size_t nq = 0;
  double *t1s = nullptr;  // type-1 head -> tail state: p, q, I planes of nq doubles
  int *tq1 = nullptr;     // type-1 tail queues (per class, indexed like the units)
  unsigned long long *qf = nullptr;
  unsigned long long *qopen = nullptr, *qfbf = nullptr;  // speculative type-1 fallback
  double *qfb = nullptr;
  double *udir = nullptr;  // per type-1 unit: its pair direction (4 values)
  int *fbq1 = nullptr, *fbc = nullptr;
  unsigned char *stat = nullptr;
  unsigned char *t1fb = nullptr, *t2fb = nullptr;  // per triplet: has large-grid work (type 1 / type 2)
  ClassMap cm_head, cm_tail, cm_fb1, cm_fb2;      // fused launches
  int max_lab = 0, max_lmb = 0;
  int tcount2[kMaxLMB + 1] = {0};  // triplets per type-2 fallback class
  int max_cls2 = 0;
  int *tclaim = nullptr;
  unsigned char *fbf2 = nullptr;  // per T entry: the speculative fallback failed on the large grid
  const unsigned char *tcls2 = nullptr;
  const int *tgf = nullptr;
  int max_lbs = 0, max_L = 0;  // largest basis l and ECP local label of the job
  const T2Group *grp = nullptr;
  const int *goff = nullptr;
  double *Tout = nullptr;
  unsigned char *tf = nullptr;
  int *gflag = nullptr, *fbq2 = nullptr;
  int *flist = nullptr;  // per group: its failed T entries (gflag[g] of them) at eoff
  int *tq2 = nullptr;    // type-2 entries left open by the head kernel
  double *fbT = nullptr;
  const int2 *fit = nullptr, *ait = nullptr;
  AsmArgs aa{};
  AsmArgs aa4{};  // packed type-1 assembly of classes la + lb <= kAsmPackLab
  unsigned int *zero_base = nullptr;
  size_t zero_words = 0;
  double *blk = nullptr;
  size_t asm1_lds = 0, asm2_lds = 0, asm1_lds4 = 0, asm2_lds4 = 0;
  AsmArgs aa2p{};  // packed type-2 assembly of classes la + lb <= kAsmPackLab
  int tcls_begin[ECP_MAX_LAB + 2] = {0};  // triplets of each class lambda = la + lb in tcls
  const int *tcls = nullptr;
  // algorithmic FLOPs per evaluation (the per-unit models of the reference's
  // eager algorithm): type-1 radial work per class, type-2 radial work
  double flops_t1[ECP_MAX_LAB + 1] = {0};
  double flops_t2 = 0;
  double model[ECP_FL_N] = {0};  // the model's terms known before evaluation (ecp_model_flops)
};

}  // namespace

constexpr int kWorkers = 4;
constexpr int kSpanGraphs = 8;  // graph copies with their own quadrature event pair

struct ecp_dev {
  hipStream_t stream = nullptr;
  hipStream_t workers[kWorkers] = {};
  hipEvent_t fork = nullptr, t2done = nullptr, t1done = nullptr, angdone = nullptr, join[kWorkers] = {};
  hipEvent_t h1done = nullptr;  // type-1 head finished (speculative fallback)
  hipEvent_t h2done = nullptr;  // type-2 head finished (speculative fallback)
  // the type-1 small-grid quadrature (head and tail launches), recorded on
  // its stream: event pair k belongs to captured graph k (replays cycle over
  // the graphs, so the last kSpanGraphs replays can be read); the eager
  // evaluation uses pair 0
  hipEvent_t span[kSpanGraphs][2] = {};
  int span_k = 0;
  bool span_valid = false;
  bool capturing = false;  // enqueue runs under stream capture
  bool last_eager = false;  // the last execute ran no replay
  long long replays = 0;    // replays since the graphs were captured
  int last_replays = 0;     // replays of the last execute
  DevTables T{};
  void *static_mem = nullptr;
  char *arena = nullptr;
  size_t arena_cap = 0;
  char *pin_in = nullptr;
  size_t pin_in_cap = 0;
  char *pin_out = nullptr;
  size_t pin_out_cap = 0;
  double *dfac_dev = nullptr, *fac_dev = nullptr;
  int lds_max = 64 * 1024;  // LDS per workgroup of the device (160 KiB on gfx950)
  hipEvent_t ev[8] = {};
  float last_ms[8] = {};
  int nev = 0;
  std::vector<int> tmpl, tmpl_off, tmpl_cnt;
  int *tmpl_dev = nullptr;
  Prepared P;
  JobLayout L;
  ecp_tuning tune{8, 8, 100000, 8192, 2, 1, 15, 2048, 0, 1, 0};
  bool planned = false, prepared = false, eager_done = false;
  hipGraph_t graph[kSpanGraphs] = {};
  hipGraphExec_t gexec[kSpanGraphs] = {};
  int ngraph = 0;
  // per-launch profiling (eager runs only)
  bool profiling = false;
  std::vector<hipEvent_t> pev;
  std::vector<int> pfam;
  int npev = 0;
  // density contraction: index tables, densities, partial sums, results
  char *con_mem = nullptr;
  size_t con_cap = 0;
  hipEvent_t con_ev[2] = {};
  std::vector<char> con_idx;  // index tables last uploaded to con_mem at con_idx_at
  size_t con_idx_at = 0;
  double *fbf = nullptr;      // the profile's fallback FLOP sums
  void *scratch = nullptr;    // this device's no-allocation scratch (scratch_setup)
};

namespace {
// kernel families reported by the profiler
enum { FAM_TABLES = 0, FAM_TYPE1, FAM_TYPE2, FAM_FB1, FAM_FB2, FAM_ASM, FAM_T1TAIL, FAM_N };
const char *kFamName[FAM_N] = {"tables", "type1", "type2", "fallback1", "fallback2", "assemble", "type1_tail"};

// records an event pair around one launch on stream s when profiling
struct Span {
  ecp_dev *d;
  hipStream_t s;
  int i;
  Span(ecp_dev *d_, hipStream_t s_, int fam) : d(d_), s(s_), i(-1) {
    if (!d->profiling) return;
    if ((size_t)(2 * d->npev + 2) > d->pev.size()) {
      hipEvent_t a, b;
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      d->pev.push_back(a);
      d->pev.push_back(b);
      d->pfam.push_back(fam);
    }
    i = d->npev++;
    d->pfam[i] = fam;
    (void)hipEventRecord(d->pev[2 * i], s);
  }
  ~Span() {
    if (i >= 0) (void)hipEventRecord(d->pev[2 * i + 1], s);
  }
};
}  // namespace

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

void reset_graph(ecp_dev *d) {
  for (int k = 0; k < d->ngraph; k++) {
    if (d->gexec[k]) (void)hipGraphExecDestroy(d->gexec[k]);
    if (d->graph[k]) (void)hipGraphDestroy(d->graph[k]);
    d->gexec[k] = nullptr;
    d->graph[k] = nullptr;
  }
  d->ngraph = 0;
  d->replays = 0;
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
  {
    int dev = 0, v = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return fail("hipGetDevice", e);
    if ((e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
      return fail("hipDeviceGetAttribute", e);
    d->lds_max = v;
  }
  if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) return fail("hipStreamCreate", e);
  for (int i = 0; i < kWorkers; i++) {
    if ((e = hipStreamCreateWithFlags(&d->workers[i], hipStreamNonBlocking)) != hipSuccess)
      return fail("hipStreamCreate", e);
    if ((e = hipEventCreateWithFlags(&d->join[i], hipEventDisableTiming)) != hipSuccess)
      return fail("hipEventCreate", e);
  }
  if ((e = hipEventCreateWithFlags(&d->fork, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = hipEventCreateWithFlags(&d->t2done, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = hipEventCreateWithFlags(&d->t1done, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = hipEventCreateWithFlags(&d->angdone, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = hipEventCreateWithFlags(&d->h1done, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  if ((e = hipEventCreateWithFlags(&d->h2done, hipEventDisableTiming)) != hipSuccess) return fail("hipEventCreate", e);
  for (int k = 0; k < kSpanGraphs; k++)
    for (int i = 0; i < 2; i++)
      if ((e = hipEventCreate(&d->span[k][i])) != hipSuccess) return fail("hipEventCreate", e);
  for (int i = 0; i < 8; i++)
    if ((e = hipEventCreate(&d->ev[i])) != hipSuccess) return fail("hipEventCreate", e);
  for (int i = 0; i < 2; i++)
    if ((e = hipEventCreate(&d->con_ev[i])) != hipSuccess) return fail("hipEventCreate", e);
  if (scratch_setup(&d->scratch, err, errlen)) {
    ecp_dev_destroy(d);
    return nullptr;
  }
  build_templates(d);
  const int so = ECP_SMALL_ORDER, lo = st->lorder;
  const size_t nbess = (size_t)(ECP_BESSEL_N + 1) * ECP_BESSEL_STRIDE;
  const size_t npoly = (size_t)ECP_DIM_C(st->tmDim) * ECP_DIM_L(st->tmDim);
  const size_t nom = (size_t)ECP_DIM_L(st->omL) * ECP_DIM_L(st->omLambda) * ECP_DIM_C(st->omAlpha);
  const size_t nijk = 3 * (size_t)ECP_DIM_C(st->ijkDim - 1);
  const size_t nidx = (size_t)st->ijkDim * st->ijkDim * st->ijkDim;
  const size_t nfac = 2 * (size_t)st->tmDim + 6;
  std::vector<double> binom(16 * 16, 0.0);
  {
    double f[16];
    f[0] = 1.0;
    for (int i = 1; i < 16; i++) f[i] = f[i - 1] * i;
    for (int n = 0; n < 16; n++)
      for (int k = 0; k <= n; k++) binom[n * 16 + k] = f[n] / (f[n - k] * f[k]);
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
  reset_graph(d);
  if (d->static_mem) (void)hipFree(d->static_mem);
  if (d->arena) (void)hipFree(d->arena);
  if (d->pin_in) (void)hipHostFree(d->pin_in);
  if (d->pin_out) (void)hipHostFree(d->pin_out);
  if (d->fbf) (void)hipFree(d->fbf);
  for (int i = 0; i < 8; i++)
    if (d->ev[i]) (void)hipEventDestroy(d->ev[i]);
  for (int i = 0; i < kWorkers; i++) {
    if (d->workers[i]) (void)hipStreamSynchronize(d->workers[i]);
    if (d->workers[i]) (void)hipStreamDestroy(d->workers[i]);
    if (d->join[i]) (void)hipEventDestroy(d->join[i]);
  }
  if (d->fork) (void)hipEventDestroy(d->fork);
  if (d->t2done) (void)hipEventDestroy(d->t2done);
  if (d->t1done) (void)hipEventDestroy(d->t1done);
  if (d->angdone) (void)hipEventDestroy(d->angdone);
  if (d->h1done) (void)hipEventDestroy(d->h1done);
  if (d->h2done) (void)hipEventDestroy(d->h2done);
  for (int k = 0; k < kSpanGraphs; k++)
    for (int i = 0; i < 2; i++)
      if (d->span[k][i]) (void)hipEventDestroy(d->span[k][i]);
  for (auto e : d->pev) (void)hipEventDestroy(e);
  if (d->con_mem) (void)hipFree(d->con_mem);
  for (int i = 0; i < 2; i++)
    if (d->con_ev[i]) (void)hipEventDestroy(d->con_ev[i]);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

extern "C" int ecp_dev_last_timing(ecp_dev *d, double *ms, int n) {
  int k = std::min(n, d->nev);
  for (int i = 0; i < k; i++) ms[i] = d->last_ms[i];
  return k;
}

namespace {

// Type-1 small grid of one class: the head kernel, then -- for the pairs
// still open after `head` levels -- the wave-cooperative tail when the class
// is small (latency-bound: few waves on the device) or the compacted
// lane-serial tail when it is large (throughput-bound).
// Type-1 small grid of one class: without a tail (head >= the level count)
// one lane-serial kernel; otherwise the head kernel, then for the pairs still
// open the wave-cooperative tail when the class is small (latency-bound: few
// waves on the device) or the compacted lane-serial tail when it is large.
// type-2 small grid: the lane-per-entry kernel, split into head and
// compacted tail when t2_head is below the level count
// PS93 levels of the type-2 head (t2_head; 0: automatic -- 6 when the small
// grid is on the critical chain (fused launches), none in large per-class
// jobs)
int t2_head_levels(const ecp_dev *d, bool fused) {
  return d->tune.t2_head > 0 ? d->tune.t2_head : (fused ? 6 : 15);
}

// spec: the head enlists every entry it leaves open for the fallback, which
// can start at `mid` (recorded after the head) beside the tail
void launch_type2(ecp_dev *d, hipStream_t s, bool fused, bool spec = false, hipEvent_t mid = nullptr) {
  const Prepared &P = d->P;
  if (P.ntot2 <= 0) return;
  T2Args A{P.grp, P.goff, P.ngroups, P.ntot2, d->tmpl_dev, P.Tout, P.tf, P.gflag, P.fbq2, P.fbc + kFb2Cnt, P.nt,
           P.t2fb, P.flist, P.tq2, P.fbc + 48, P.tclaim, P.tcls2, spec ? 1 : 0, P.fbf2};
  const int head = t2_head_levels(d, fused);
  {
    Span sp(d, s, FAM_TYPE2);
    k_type2<<<(P.ntot2 + 127) / 128, 128, 0, s>>>(d->T, P.J, A, head);
  }
  if (mid) (void)hipEventRecord(mid, s);
  if (head < d->T.s93.nlevels) {
    Span sp(d, s, FAM_TYPE2);
    k_type2_tail<<<std::min((P.ntot2 + 15) / 16, d->tune.t2_tail_grid), 64, 0, s>>>(d->T, P.J, A);
  }
}

// type-2 assembly: the register-cached variant fits basis l <= 3 with
// projectors l <= 3 (NB = 4 lambda2 values, NM = 7 m values) or basis
// l <= 4 with projectors l <= 4; larger jobs read the tables directly.  The
// cache costs occupancy: it pays where the assembly is on the critical path
// of a small job (fused launches), not in large throughput-bound ones.
// pack: the triplets of classes la + lb <= kAsmPackLab (the first of the
// class-ordered list) kAsmPack to a wave, the others one per workgroup
template <int TPB>
void launch_asm2_tpb(ecp_dev *d, hipStream_t s, int filter, bool cache, bool spec, bool pack) {
  const Prepared &P = d->P;
  AsmArgs aa = P.aa;
  aa.fbf2 = spec ? P.fbf2 : nullptr;
  const int m = pack ? P.tcls_begin[kAsmPackLab + 1] : 0;
  if (m > 0) {
    AsmArgs a2 = P.aa2p;
    a2.fbf2 = aa.fbf2;
    const int g = (m + kAsmPack - 1) / kAsmPack;
    // the 2 l + 1 angular factors loaded together where l < 4 (148
    // registers, three waves per SIMD, against 84 and five read pair by
    // pair: the uneven batch 0.7 ms faster)
    if (P.max_L <= 4) k_asm2p<7><<<g, kAsmThreads, P.asm2_lds4, s>>>(d->T, P.J, a2, P.stat, P.tcls, m, P.t2fb, filter);
    else k_asm2p<1><<<g, kAsmThreads, P.asm2_lds4, s>>>(d->T, P.J, a2, P.stat, P.tcls, m, P.t2fb, filter);
  }
  const int n = P.nt - m;
  if (n <= 0) return;
  const int *tl = m > 0 ? P.tcls + m : nullptr;
  if (!cache && P.max_L <= 4)
    k_asm2<0, 7, TPB><<<n, TPB, P.asm2_lds, s>>>(d->T, P.J, aa, P.stat, P.t2fb, filter, tl);
  else if (!cache)
    k_asm2<0, 1, TPB><<<n, TPB, P.asm2_lds, s>>>(d->T, P.J, aa, P.stat, P.t2fb, filter, tl);
  else if (P.max_lbs <= 3 && P.max_L <= 4)
    k_asm2<4, 7, TPB><<<n, TPB, P.asm2_lds, s>>>(d->T, P.J, aa, P.stat, P.t2fb, filter, tl);
  else if (P.max_lbs <= 4 && P.max_L <= 5)
    k_asm2<5, 9, TPB><<<n, TPB, P.asm2_lds, s>>>(d->T, P.J, aa, P.stat, P.t2fb, filter, tl);
  else
    k_asm2<0, 1, TPB><<<n, TPB, P.asm2_lds, s>>>(d->T, P.J, aa, P.stat, P.t2fb, filter, tl);
}

// type-1 assembly of the triplets tcls[b .. e) (class-ordered), one per
// wave, or with pack those of classes la + lb <= kAsmPackLab kAsmPack to a
// wave (one more launch: it pays in large per-class launches, not on the
// critical chain of fused ones)
// cls >= 0: every triplet of the range is of class la + lb = cls (per-class
// launches), which bounds the monomial degrees the kernel holds in registers
// spec: the type-1 fallback ran speculatively (T1Set::qfb), the assembly
// picks its results where the small grid failed
void launch_asm1(ecp_dev *d, hipStream_t s, int b, int e, int filter, bool pack, int cls = -1, bool spec = false) {
  const Prepared &P = d->P;
  AsmArgs aa = P.aa;
  if (spec) {
    aa.qfail = P.qf;
    aa.qfbf = P.qfbf;
    aa.qfb = P.qfb;
  }
  const int m = pack ? std::min(std::max(b, P.tcls_begin[kAsmPackLab + 1]), e) : b;
  if (m > b) {
    AsmArgs a4 = P.aa4;
    if (spec) {
      a4.qfail = P.qf;
      a4.qfbf = P.qfbf;
      a4.qfb = P.qfb;
    }
    k_asm1<kAsmPack, kAsmPackLab><<<(m - b + kAsmPack - 1) / kAsmPack, kAsmThreads, P.asm1_lds4, s>>>(
        d->T, P.J, a4, P.stat, P.tcls + b, m - b, P.t1fb, filter);
  }
  if (e <= m) return;
  const dim3 g(e - m);
  switch (cls) {
#define ASM1(D)                                                                                               \
  case D:                                                                                                   \
    k_asm1<1, D><<<g, kAsmThreads, P.asm1_lds, s>>>(d->T, P.J, aa, P.stat, P.tcls + m, e - m, P.t1fb, filter); \
    break;
    ASM1(0) ASM1(1) ASM1(2) ASM1(3) ASM1(4) ASM1(5)
#undef ASM1
    default:
      // classes above 5 and mixed-class launches: rows up to degree 6 in
      // registers (the uneven batch's class-6 chain 0.7 ms slower with
      // degree 6 in the general loop)
      k_asm1<1, 6><<<g, kAsmThreads, P.asm1_lds, s>>>(d->T, P.J, aa, P.stat, P.tcls + m, e - m, P.t1fb, filter);
  }
}

// The assembly of the triplets with large-grid work (filter 2) comes last
// and covers few triplets: with fused launches it runs 4 waves per triplet,
// so that each triplet's elements finish sooner.
// spec: the fallback ran speculatively; a triplet failed where an entry
// failed on both grids
// The small classes go four triplets to a wave (they are about half of the
// triplets: the medium cluster 2.5 % and the uneven batch 0.5 % faster),
// except in the fused sequence's last assembly, which covers the few
// triplets with fallback work at four waves each.
void launch_asm2(ecp_dev *d, hipStream_t s, int filter, bool fused, bool spec = false) {
  const int mode = d->tune.asm2_cache;
  const bool cache = mode == 1 || (mode == 2 && fused);
  if (fused && filter == 2)
    launch_asm2_tpb<4 * kAsmThreads>(d, s, filter, cache, spec, false);
  else
    launch_asm2_tpb<kAsmThreads>(d, s, filter, cache, spec, true);
}

template <int LAB>
void launch_type1(ecp_dev *d, hipStream_t s, const DevTables &T, const DevJob &J, const T1Unit *u, int n, double *q,
                  unsigned long long *qf, int *fbq, int *fbc, double *st, size_t plane, int *tq, int *tc,
                  unsigned char *ff) {
  if (n <= 0) return;
  const bool coop = n <= d->tune.t1_coop_max;
  const int head = coop ? d->tune.t1_head_coop : d->tune.t1_head_serial;
  const int grid = (n + 63) / 64;
  if (head >= T.s93.nlevels) {
    Span sp(d, s, FAM_TYPE1);
    k_type1<LAB, 0><<<grid, 64, 0, s>>>(T, J, u, n, q, qf, fbq, fbc, head, st, plane, tq, tc, ff);
    return;
  }
  {
    Span sp(d, s, FAM_TYPE1);
    k_type1<LAB, 1><<<grid, 64, 0, s>>>(T, J, u, n, q, qf, fbq, fbc, head, st, plane, tq, tc, ff);
  }
  Span sp(d, s, FAM_T1TAIL);
  if (coop)
    k_t1tail<LAB><<<std::min(n, d->tune.t1_tail_grid), 64, 0, s>>>(T, J, u, q, qf, fbq, fbc, head, st, plane, tq,
                                                                    tc, ff);
  else
    k_type1<LAB, 2><<<std::min(grid, d->tune.t1_tail_grid), 64, 0, s>>>(T, J, u, n, q, qf, fbq, fbc, head, st,
                                                                         plane, tq, tc, ff);
}

template <int LAB>
void launch_fb1(ecp_dev *d, hipStream_t s, const DevTables &T, const DevJob &J, const T1Unit *u, int n, double *q,
                const unsigned long long *qf, const int *fbq, const int *fbc, unsigned char *st) {
  if (n > 0) {
    Span sp(d, s, FAM_FB1);
    const int blocks = std::min(n, 4096);
    k_fb1<LAB><<<blocks, 64, 0, s>>>(T, J, u, q, qf, fbq, fbc, st);
  }
}

// the type-2 fallback classes of cm in one grid; like the type-1 kernels,
// three tiers bound the Bessel orders the grid holds: up to 7, 11 and 12
// pack: several pairs per wave for triplets with few failed entries
// (classes up to 7).  Less work in all, but each wave longer: it pays in
// the throughput-bound per-class sequence of large jobs (the uneven batch
// 1.4 % faster) and not on the fused sequence's critical chain (the medium
// cluster 1.5 % slower).
void launch_fb2_fused(ecp_dev *d, hipStream_t s, const T2Set &S2, const ClassMap &cm, bool pack) {
  const Prepared &P = d->P;
  const int nb = cm.bbeg[cm.n];
  if (P.max_cls2 > 11) k_fb2_fused<12, false><<<nb, 64, 0, s>>>(d->T, P.J, S2, cm);
  else if (P.max_cls2 > 7) k_fb2_fused<11, false><<<nb, 64, 0, s>>>(d->T, P.J, S2, cm);
  else if (pack) k_fb2_fused<7, true><<<nb, 64, 0, s>>>(d->T, P.J, S2, cm);
  else k_fb2_fused<7, false><<<nb, 64, 0, s>>>(d->T, P.J, S2, cm);
}

// waves per (triplet, pair) of the type-2 fallback (fb2_split; 0: automatic,
// 3 for jobs of up to kFb2SplitMaxItems (triplet, pair) items, where the
// fallback's latency counts, 1 for larger, throughput-bound ones)
constexpr long kFb2SplitMaxItems = 1000000;

int fb2_split(const ecp_dev *d, bool fused) {
  (void)fused;
  const int v = d->tune.fb2_split;
  if (v > 0) return std::min(v, 8);
  long items = 0;
  for (int c = 0; c <= kMaxLMB; c++) items += (long)d->P.tcount2[c] * d->P.maxPairs;
  return items <= kFb2SplitMaxItems ? 3 : 1;
}

// The kernel sequence of one evaluation, on resident inputs, one launch per
// class: tables on the main stream, then the type-2 chain and the type-1
// classes, each class as small grid -> large-grid fallback -> assembly, on
// forked streams (two layouts, see enqueue).  With evs, stage boundaries are
// recorded on the main stream: start, tables done, end, end (assembly
// overlaps the quadratures).
void enqueue_fused(ecp_dev *d, hipStream_t s, hipEvent_t *evs);

void launch_tables(ecp_dev *d, hipStream_t s) {
  const Prepared &P = d->P;
  TabArgs A{};
  A.zero_base = P.zero_base;
  A.zero_words = P.zero_words;
  A.nzb = (int)std::max<size_t>(1, std::min<size_t>((P.zero_words + 127) / 128, 2048));
  A.rows = P.max_L + 1;
  A.nub = P.ncent * A.rows * kUtabChunks;
  A.nfit = P.nfit;
  A.fit = P.fit;
  A.units = P.u1;
  A.nunits = P.cls_begin[ECP_MAX_LAB + 1];
  A.udir = P.udir;
  k_tables<<<A.nzb + A.nub + A.nfit + (A.nunits + 127) / 128, 128, 0, s>>>(d->T, P.J, A);
}

// the angular tables on stream s, concurrent with the quadratures; the
// type-2 assembly waits for d->angdone
void launch_ang(ecp_dev *d, hipStream_t s) {
  const Prepared &P = d->P;
  if (P.nait) k_ang<<<P.nait, 128, 0, s>>>(d->T, P.J, P.ait, d->fac_dev, d->dfac_dev);
  (void)hipEventRecord(d->angdone, s);
}

// fuse = 2 (automatic): fused launches for jobs of up to this many type-1
// primitive pairs (latency-bound: per-class launches hold few waves each)
constexpr int kFuseAutoMaxUnits = 150000;

// workgroup ranges of the classes with work; blocks(c) = workgroups of class c
template <typename F>
ClassMap class_map(int ncls, const int *count, F blocks) {
  ClassMap cm{};
  int b = 0;
  for (int c = 0; c < ncls; c++) {
    if (count[c] <= 0) continue;
    const int nb = blocks(c);
    if (nb <= 0) continue;
    cm.cls[cm.n] = c;
    cm.bbeg[cm.n] = b;
    b += nb;
    cm.n++;
  }
  cm.bbeg[cm.n] = b;
  return cm;
}

// one type-1 class of the per-class sequence: small grid, fallback, assembly
template <int c>
void enqueue_t1_class_t(ecp_dev *d, hipStream_t ws) {
  const Prepared &P = d->P;
  launch_type1<c>(d, ws, d->T, P.J, P.u1 + P.cls_begin[c], P.cls_count[c], P.q, P.qf + P.cls_begin[c],
                  P.fbq1 + P.cls_begin[c], P.fbc + c, P.t1s, P.nq, P.tq1 + P.cls_begin[c], P.fbc + 32 + c, P.t1fb);
  launch_fb1<c>(d, ws, d->T, P.J, P.u1 + P.cls_begin[c], P.cls_count[c], P.q, P.qf + P.cls_begin[c],
                P.fbq1 + P.cls_begin[c], P.fbc + c, P.stat);
  if (P.tcls_begin[c + 1] > P.tcls_begin[c]) {
    Span sp(d, ws, FAM_ASM);
    launch_asm1(d, ws, P.tcls_begin[c], P.tcls_begin[c + 1], 0, true, c);
  }
}

void enqueue_t1_class(ecp_dev *d, hipStream_t ws, int c) {
  switch (c) {
#define T1(c) \
  case c:     \
    enqueue_t1_class_t<c>(d, ws); \
    break;
    T1(0) T1(1) T1(2) T1(3) T1(4) T1(5) T1(6) T1(7) T1(8) T1(9) T1(10) T1(11) T1(12)
#undef T1
  }
}

void enqueue(ecp_dev *d, hipStream_t s, hipEvent_t *evs) {
  const int units = d->P.cls_begin[ECP_MAX_LAB + 1];
  const bool fused = d->tune.fuse == 1 || (d->tune.fuse == 2 && units <= kFuseAutoMaxUnits);
  // point-major type-2 rows in the throughput-bound per-class sequence (the
  // uneven batch 1.8 % faster); row-major in the fused one, where the
  // point-major tables cost the medium cluster 2.7 %
  d->P.J.ptmajor = fused ? 0 : 1;
  if (fused) {
    enqueue_fused(d, s, evs);
    return;
  }
  const Prepared &P = d->P;
  const DevTables &T = d->T;
  const DevJob &J = P.J;
  hipStream_t *w = d->workers;
  d->span_valid = false;  // type-1 classes run on several streams
  if (evs) (void)hipEventRecord(evs[0], s);
  {
    Span sp(d, s, FAM_TABLES);
    launch_tables(d, s);
  }
  if (evs) (void)hipEventRecord(evs[1], s);
  (void)hipEventRecord(d->fork, s);
  for (int i = 0; i < kWorkers; i++) (void)hipStreamWaitEvent(w[i], d->fork, 0);
  // Two layouts.  Jobs whose type-1 classes stay at la + lb <= 6: four
  // streams, one per hardware queue the device gives a process -- the
  // type-2 chain (small grid, fallbacks, assembly of the triplets with
  // fallback work) on the main stream, the assembly of the other triplets
  // on the third worker, the type-1 classes on the first two.  Jobs with
  // larger classes, where type 1 carries more of the work: the type-2 small
  // grid and the assembly without fallback work on the fourth worker, the
  // fallbacks and the rest of the assembly on the main stream, the type-1
  // classes on the first three workers.
  const bool four = P.max_lab <= 6;
  const int nw1 = four ? 2 : kWorkers - 1;
  hipStream_t w2 = four ? s : w[kWorkers - 1];
  hipStream_t wa = four ? w[2] : w2;  // the assembly without fallback work
  {
    Span sp(d, s, FAM_TABLES);
    launch_ang(d, s);
  }
  if (P.ntot2 > 0) {
    launch_type2(d, w2, false);
    (void)hipEventRecord(d->t2done, w2);
    // the angular tables are on the main stream: behind the small grid in
    // the four-stream layout, a wait otherwise
    (void)hipStreamWaitEvent(wa, four ? d->t2done : d->angdone, 0);
    {
      Span sp(d, wa, FAM_ASM);
      launch_asm2(d, wa, 1, false);
    }
    if (!four) (void)hipStreamWaitEvent(s, d->t2done, 0);
    // every fallback class in one persistent grid: the classes' items run
    // side by side instead of one class after another on this stream
    {
      const int nsplit = fb2_split(d, false);
      T2Set S2{P.grp, d->tmpl_dev, P.tf, P.fbq2, P.fbc, P.fbT, P.stat, P.nt, P.gflag, P.flist, P.tgf, nsplit};
      const ClassMap cm2 = class_map(kMaxLMB + 1, P.tcount2, [&](int c) {
        return (int)std::min<long>((long)P.tcount2[c] * P.maxPairs * nsplit, 8192);
      });
      if (cm2.n > 0) {
        Span sp(d, s, FAM_FB2);
        launch_fb2_fused(d, s, S2, cm2, true);
      }
    }
    {
      Span sp(d, s, FAM_ASM);
      launch_asm2(d, s, 2, false);
    }
  }
  // type-1 classes, each followed by its fallback and assembly on one of the
  // type-1 workers: the costliest class first, each on the worker with the
  // least work so far (the cost of a class: its algorithmic small-grid
  // FLOPs), so that no large class starts last and runs alone at the end
  int order[ECP_MAX_LAB + 1], nord = 0;
  for (int c = 0; c <= ECP_MAX_LAB; c++)
    if (P.cls_count[c] > 0) order[nord++] = c;
  std::stable_sort(order, order + nord, [&](int a, int b) { return P.flops_t1[a] > P.flops_t1[b]; });
  double load[kWorkers - 1] = {};
  for (int i = 0; i < nord; i++) {
    int k = 0;
    for (int j = 1; j < nw1; j++)
      if (load[j] < load[k]) k = j;
    load[k] += P.flops_t1[order[i]];
    enqueue_t1_class(d, w[k], order[i]);
  }
  for (int i = 0; i < kWorkers; i++) {
    (void)hipEventRecord(d->join[i], w[i]);
    (void)hipStreamWaitEvent(s, d->join[i], 0);
  }
  if (evs) {
    (void)hipEventRecord(evs[2], s);
    (void)hipEventRecord(evs[3], s);
  }
}


// The fused sequence: tables on the main stream, then two chains.  Type 1:
// head (all classes) -> tail -> [assembly of triplets without large-grid
// work] while, on a second stream, large-grid fallback -> assembly of the
// rest.  Type 2 likewise: small grid -> [assembly without fallback] while
// fallback -> assembly of the rest.  evs: stage boundaries as in enqueue.
void enqueue_fused(ecp_dev *d, hipStream_t s, hipEvent_t *evs) {
  const Prepared &P = d->P;
  const DevTables &T = d->T;
  const DevJob &J = P.J;
  hipStream_t *w = d->workers;
  d->span_valid = false;
  if (evs) (void)hipEventRecord(evs[0], s);
  {
    Span sp(d, s, FAM_TABLES);
    launch_tables(d, s);
  }
  if (evs) (void)hipEventRecord(evs[1], s);
  (void)hipEventRecord(d->fork, s);
  for (int i = 0; i < kWorkers; i++) (void)hipStreamWaitEvent(w[i], d->fork, 0);
  const ecp_tuning &tu = d->tune;
  // ---- type 1 ----
  T1Set S{};
  S.units = P.u1;
  S.q = P.q;
  S.qf = P.qf;
  S.fbq = P.fbq1;
  S.fbc = P.fbc;
  S.tq = P.tq1;
  S.st = P.t1s;
  S.plane = P.nq;
  S.stat = P.stat;
  S.fbflag = P.t1fb;
  // cooperative tail for classes up to t1_coop_max pairs, none for larger
  // ones (throughput-bound: the lane-serial head runs every level)
  int tailcnt[ECP_MAX_LAB + 1];
  bool tail = false;
  for (int c = 0; c <= ECP_MAX_LAB; c++) {
    S.ubeg[c] = P.cls_begin[c];
    S.ucnt[c] = P.cls_count[c];
    const bool coop = P.cls_count[c] <= tu.t1_coop_max && tu.t1_head_coop < T.s93.nlevels;
    S.head[c] = coop ? tu.t1_head_coop : T.s93.nlevels;
    tailcnt[c] = coop ? P.cls_count[c] : 0;
    tail = tail || (coop && P.cls_count[c] > 0);
  }
  // the fused kernels hold the registers of their largest class: three
  // instantiations, classes up to 6, 10 and 12
  const int tier1 = P.max_lab > 10 ? 2 : P.max_lab > 6 ? 1 : 0;
  // speculative fallback: with a tail, the large-grid fallback starts after
  // the head on every entry the head left open, beside the tail, instead of
  // after the tail on the failed ones (most of the entries open after the
  // head fail; the tail only decides which results the assembly takes)
  const bool spec = tail && tu.t1_spec != 0;
  if (spec) {
    S.qopen = P.qopen;
    S.qfb = P.qfb;
    S.qfbf = P.qfbf;
  }
  const ClassMap cmh = class_map(ECP_MAX_LAB + 1, P.cls_count, [&](int c) { return (P.cls_count[c] + 63) / 64; });
  if (cmh.n > 0) {
    // external event nodes when captured, so that replays record them too
    (void)hipEventRecordWithFlags(d->span[d->span_k][0], w[0], d->capturing ? hipEventRecordExternal : 0);
    {
      Span sp(d, w[0], FAM_TYPE1);
      if (tier1 == 2) k_type1_fused<12, 1><<<cmh.bbeg[cmh.n], 64, 0, w[0]>>>(T, J, S, cmh);
      else if (tier1 == 1) k_type1_fused<10, 1><<<cmh.bbeg[cmh.n], 64, 0, w[0]>>>(T, J, S, cmh);
      else k_type1_fused<6, 1><<<cmh.bbeg[cmh.n], 64, 0, w[0]>>>(T, J, S, cmh);
    }
    if (spec) (void)hipEventRecord(d->h1done, w[0]);
    if (tail) {
      const ClassMap cmt = class_map(ECP_MAX_LAB + 1, tailcnt,
                                     [&](int c) { return std::min(tailcnt[c], tu.t1_tail_grid); });
      Span sp(d, w[0], FAM_T1TAIL);
      if (tier1 == 2) k_t1tail_fused<12><<<cmt.bbeg[cmt.n], 64, 0, w[0]>>>(T, J, S, cmt);
      else if (tier1 == 1) k_t1tail_fused<10><<<cmt.bbeg[cmt.n], 64, 0, w[0]>>>(T, J, S, cmt);
      else k_t1tail_fused<6><<<cmt.bbeg[cmt.n], 64, 0, w[0]>>>(T, J, S, cmt);
    }
    (void)hipEventRecordWithFlags(d->span[d->span_k][1], w[0], d->capturing ? hipEventRecordExternal : 0);
    d->span_valid = true;
    // the fallback is captured before the assembly beside it: a graph
    // continues a node's queue with its first dependent, and the fallback
    // is the critical one
    (void)hipEventRecord(d->t1done, w[0]);
    (void)hipStreamWaitEvent(w[1], spec ? d->h1done : d->t1done, 0);
    const ClassMap cmf = class_map(ECP_MAX_LAB + 1, P.cls_count, [&](int c) { return std::min(P.cls_count[c], 4096); });
    {
      Span sp(d, w[1], FAM_FB1);
      if (tier1 == 2 && spec) k_fb1_fused<12, true><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
      else if (tier1 == 2) k_fb1_fused<12, false><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
      else if (tier1 == 1 && spec) k_fb1_fused<10, true><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
      else if (tier1 == 1) k_fb1_fused<10, false><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
      else if (spec) k_fb1_fused<6, true><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
      else k_fb1_fused<6, false><<<cmf.bbeg[cmf.n], 64, 0, w[1]>>>(T, J, S, cmf);
    }
    {
      Span sp(d, w[0], FAM_ASM);
      launch_asm1(d, w[0], 0, P.nt, 1, false);
    }
    // the assembly of the triplets with large-grid work reads the tail's
    // decisions and the fallback's results
    if (spec) (void)hipStreamWaitEvent(w[1], d->t1done, 0);
    {
      Span sp(d, w[1], FAM_ASM);
      launch_asm1(d, w[1], 0, P.nt, 2, false, -1, spec);
    }
  }
  // ---- type 2 ----
  {
    Span sp(d, s, FAM_TABLES);
    launch_ang(d, s);
  }
  if (P.ntot2 > 0) {
    // likewise the fallback first (it needs no angular tables), then the
    // assembly beside it, then the assembly that waits for the fallback
    // t2_spec: the fallback starts after the head, beside the tail, on every
    // entry the head left open (most of them fail); the assembly then takes
    // the fallback's result where the tail failed too
    const bool spec2 = tu.t2_spec != 0 && t2_head_levels(d, true) < T.s93.nlevels;
    launch_type2(d, w[2], true, spec2, spec2 ? d->h2done : nullptr);
    (void)hipEventRecord(d->t2done, w[2]);
    (void)hipStreamWaitEvent(w[3], spec2 ? d->h2done : d->t2done, 0);
    const int nsplit = fb2_split(d, true);
    T2Set S2{P.grp, d->tmpl_dev, P.tf, P.fbq2, P.fbc, P.fbT, P.stat, P.nt, P.gflag, P.flist, P.tgf, nsplit};
    S2.fbf2 = spec2 ? P.fbf2 : nullptr;
    const ClassMap cm2 = class_map(kMaxLMB + 1, P.tcount2, [&](int c) {
      return (int)std::min<long>((long)P.tcount2[c] * P.maxPairs * nsplit, 8192);
    });
    if (cm2.n > 0) {
      Span sp(d, w[3], FAM_FB2);
      launch_fb2_fused(d, w[3], S2, cm2, false);
    }
    (void)hipStreamWaitEvent(w[2], d->angdone, 0);
    {
      Span sp(d, w[2], FAM_ASM);
      launch_asm2(d, w[2], 1, true);
    }
    if (spec2) (void)hipStreamWaitEvent(w[3], d->t2done, 0);
    (void)hipStreamWaitEvent(w[3], d->angdone, 0);
    {
      Span sp(d, w[3], FAM_ASM);
      launch_asm2(d, w[3], 2, true, spec2);
    }
  }
  for (int i = 0; i < kWorkers; i++) {
    (void)hipEventRecord(d->join[i], w[i]);
    (void)hipStreamWaitEvent(s, d->join[i], 0);
  }
  if (evs) {
    (void)hipEventRecord(evs[2], s);
    (void)hipEventRecord(evs[3], s);
  }
}

}  // namespace

// Host half of preparing a job: class grouping of the type-1 units, type-2
// entry groups, table work items, the algorithmic FLOP model, and the byte
// layout of the job's device block; the inputs are staged in pinned host
// memory.  *bytes = size of the device block the job needs.
extern "C" int ecp_dev_plan(ecp_dev *d, const ecp_job *job, size_t *bytes, char *err, size_t errlen) {
  Prepared &P = d->P;
  JobLayout &Ly = d->L;
  d->prepared = false;
  d->planned = false;
  d->eager_done = false;
  reset_graph(d);
  P = Prepared();
  Ly = JobLayout();
  const int nt = job->ntrip;
  P.nt = nt;
  P.ncent = job->ncent;
  P.nblk = job->nblk;
  if (bytes) *bytes = 0;
  if (nt == 0) {
    d->planned = true;
    return 0;
  }
  // type-1 units grouped by class lab, inside a class in triplet order
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    P.cls_count[tr.la + tr.lb] += job->shells[tr.sa].K * job->shells[tr.sb].K;
  }
  for (int c = 0; c <= ECP_MAX_LAB; c++) P.cls_begin[c + 1] = P.cls_begin[c] + P.cls_count[c];
  std::vector<T1Unit> units1(P.cls_begin[ECP_MAX_LAB + 1]);
  std::vector<int> t1first(nt), cls_fill(P.cls_begin, P.cls_begin + ECP_MAX_LAB + 1);
  std::vector<size_t> cls_qbase(ECP_MAX_LAB + 2, 0);
  for (int c = 0; c <= ECP_MAX_LAB; c++) cls_qbase[c + 1] = cls_qbase[c] + (size_t)P.cls_count[c] * nq_of(c);
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    const ecp_shell &Sa = job->shells[tr.sa], &Sb = job->shells[tr.sb];
    const int c = tr.la + tr.lb;
    t1first[t] = cls_fill[c];
    for (int pa = 0; pa < Sa.K; pa++)
      for (int pb = 0; pb < Sb.K; pb++) {
        const int u = cls_fill[c]++;
        units1[u] = T1Unit{t, pa, pb, (int)(cls_qbase[c] + (size_t)(u - P.cls_begin[c]) * nq_of(c))};
      }
    // per primitive pair: G (31 lambda + 74 + 4 n_Q) + 150 n_Q
    const double G = tr.ge - tr.gs;
    P.flops_t1[c] += (double)Sa.K * Sb.K * (G * (31.0 * c + 74 + 4.0 * nq_of(c)) + 150.0 * nq_of(c));
  }
  const size_t nq = cls_qbase[ECP_MAX_LAB + 1];
  std::vector<int> tcls(nt);
  {
    for (int t = 0; t < nt; t++) P.tcls_begin[job->trip[t].la + job->trip[t].lb + 1]++;
    for (int c = 0; c <= ECP_MAX_LAB; c++) P.tcls_begin[c + 1] += P.tcls_begin[c];
    std::vector<int> fill(P.tcls_begin, P.tcls_begin + ECP_MAX_LAB + 1);
    for (int t = 0; t < nt; t++) tcls[fill[job->trip[t].la + job->trip[t].lb]++] = t;
  }
  std::vector<T2Group> groups;
  std::vector<int> goff, tgfirst(nt);
  std::vector<unsigned char> tcls2(nt);
  for (int c = 0; c <= kMaxLMB; c++) P.tcount2[c] = 0;
  P.max_cls2 = 0;
  int ntot2 = 0;
  for (int t = 0; t < nt; t++) {
    const ecp_triplet &tr = job->trip[t];
    const int la = tr.la, lb = tr.lb;
    tgfirst[t] = (int)groups.size();
    {
      const int c2 = std::max(la, lb) + job->cent[tr.centre].L - 1;
      tcls2[t] = (unsigned char)c2;
      P.tcount2[c2]++;
      P.max_cls2 = std::max(P.max_cls2, c2);
    }
    for (int l = 0; l < job->cent[tr.centre].L; l++) {
      const int key = tmpl_key(la, lb, l);
      T2Group g{t, l, ntot2, d->tmpl_cnt[key], d->tmpl_off[key], std::max(la, lb) + l};
      groups.push_back(g);
      // per entry: 4 FLOPs per window point + 150 for the integrator
      P.flops_t2 += (double)g.ne * (4.0 * (tr.ge - tr.gs) + 150.0);
      goff.push_back(ntot2);
      ntot2 += g.ne;
      P.lmb_count[g.lmb]++;
    }
  }
  P.ngroups = (int)groups.size();
  P.ntot2 = ntot2;
  ecp_model_flops(job, P.model);
  for (int c = 0; c <= ECP_MAX_LAB; c++)
    if (P.cls_count[c]) P.max_lab = c;
  for (int c = 0; c <= kMaxLMB; c++)
    if (P.lmb_count[c]) P.max_lmb = c;
  goff.push_back(ntot2);
  std::vector<int2> fitems, aitems;
  for (int c = 0; c < job->ncent; c++) {
    const ecp_centre &C = job->cent[c];
    for (int run = 0; run <= job->deriv; run++)
      for (int s = 0; s < C.nsh; s++)
        if (!job->win[C.win_off + s].skip) fitems.push_back(make_int2(c, s * 4 + run));
    for (int a = 0; a < C.nat; a++) aitems.push_back(make_int2(c, a));
  }
  P.nfit = (int)fitems.size();
  P.nait = (int)aitems.size();
  P.maxPairs = job->maxK * job->maxK;
  // layout: inputs (copied), then device-only buffers; the per-run zeroed
  // counters and flags are contiguous
  size_t in_off = 0;
  auto in_take = [&](size_t b) {
    size_t o = in_off;
    in_off += align_up(b);
    return o;
  };
  Ly.i_atoms = in_take(sizeof(ecp_xyz) * job->nat);
  Ly.i_amaxl = in_take(4 * (size_t)job->nat);
  Ly.i_sh = in_take(sizeof(ecp_shell) * job->nsh);
  Ly.i_pa = in_take(8 * (size_t)job->nprim);
  Ly.i_pd = in_take(8 * (size_t)job->nprim);
  Ly.i_cent = in_take(sizeof(ecp_centre) * job->ncent);
  Ly.i_gauss = in_take(sizeof(ecp_gauss) * job->ngauss);
  Ly.i_win = in_take(sizeof(ecp_window) * job->nwin);
  Ly.i_trip = in_take(sizeof(ecp_triplet) * nt);
  Ly.i_u1 = in_take(sizeof(T1Unit) * units1.size());
  Ly.i_t1f = in_take(4 * (size_t)nt);
  Ly.i_grp = in_take(sizeof(T2Group) * groups.size());
  Ly.i_goff = in_take(4 * goff.size());
  Ly.i_tgf = in_take(4 * (size_t)nt);
  Ly.i_tcls2 = in_take((size_t)nt);
  Ly.i_fit = in_take(sizeof(int2) * fitems.size());
  Ly.i_ait = in_take(sizeof(int2) * aitems.size());
  Ly.i_tcls = in_take(4 * (size_t)nt);
  Ly.in_bytes = in_off;
  size_t dv_off = align_up(Ly.in_bytes);
  auto dv_take = [&](size_t b) {
    size_t o = dv_off;
    dv_off += align_up(b);
    return o;
  };
  Ly.d_zero = dv_off;
  Ly.d_fbc = dv_take(4 * kFbcSlots);
  Ly.d_gflag = dv_take(4 * (size_t)P.ngroups);
  Ly.d_stat = dv_take(2 * (size_t)nt);
  Ly.d_t1fb = dv_take((size_t)nt);
  Ly.d_t2fb = dv_take((size_t)nt);
  Ly.d_tclaim = dv_take(4 * (size_t)nt);
  Ly.d_zero_end = dv_off;
  Ly.d_fbf2 = dv_take((size_t)ntot2);
  Ly.d_UL = dv_take(8 * job->nUL);
  Ly.d_Ut = dv_take(8 * job->nUt);
  Ly.d_F = dv_take(8 * job->nF);
  Ly.d_ang = dv_take(8 * job->nang);
  Ly.d_q = dv_take(8 * nq);
  P.nq = nq;
  Ly.d_qf = dv_take(8 * units1.size());
  Ly.d_fbq1 = dv_take(4 * units1.size());
  Ly.d_T = dv_take(8 * (size_t)ntot2);
  Ly.d_tf = dv_take((size_t)ntot2);
  Ly.d_flist = dv_take(4 * (size_t)ntot2);
  Ly.d_tq2 = dv_take(4 * (size_t)ntot2);
  Ly.d_fbq2 = dv_take(4 * (size_t)(kMaxLMB + 1) * nt);
  Ly.d_fbT = dv_take(8 * (size_t)ntot2 * P.maxPairs);
  Ly.d_blk = dv_take(8 * job->nblk);
  Ly.d_t1s = dv_take(3 * 8 * nq);
  Ly.d_tq1 = dv_take(4 * units1.size());
  Ly.d_qfb = dv_take(8 * nq);
  Ly.d_qopen = dv_take(8 * units1.size());
  Ly.d_qfbf = dv_take(8 * units1.size());
  Ly.d_udir = dv_take(32 * units1.size());
  Ly.total = dv_off;
  Ly.FL = job->FL;
  Ly.UK = job->UK;
  Ly.angL = job->angL;
  Ly.ang_lam = job->ang_lam;
  Ly.ang_c = job->ang_c;
  Ly.deriv = job->deriv;
  {
    int LM = 0, LC = 1;
    for (int i = 0; i < job->nsh; i++) LM = std::max(LM, job->shells[i].l + job->deriv);
    for (int c = 0; c < job->ncent; c++) LC = std::max(LC, job->cent[c].L);
    P.max_lbs = LM;
    P.max_L = LC;
    AsmArgs &aa = P.aa;
    aa.pair_chunk = std::min((LM <= 4) ? 16 : 8, std::max(1, P.maxPairs));
    aa.lds_rs_stride = (2 * LM + 1) * (2 * LM + 1);
    int o = 0;
    aa.shift_lm = LM;
    const int ns = shift_tab_doubles(LM);
    aa.lds_J = o; o += ECP_DIM_IJK(LM) * ECP_DIM_C(LM);
    const int base = o;
    // type 1: H stays for the shift; the shift tables reuse rsph and Q
    aa.lds_H = o; o += ECP_DIM_C(2 * LM);
    aa.lds_S1 = o;
    aa.lds_rsph = o; o += aa.pair_chunk * aa.lds_rs_stride;
    aa.lds_Q = o; o += aa.pair_chunk * nq_of(2 * LM);
    const int end1 = std::max(o, aa.lds_S1 + ns);
    // type 2: blk stays for the shift; the shift tables reuse T
    o = base;
    aa.lds_blk = o; o += ECP_DIM_C(LM) * ECP_DIM_C(LM);
    aa.lds_S2 = o;
    aa.lds_T = o; o += (LM + LC) * (LM + LC) * (2 * LM + 1);
    const int end2 = std::max(o, aa.lds_S2 + ns);
    P.asm1_lds = (size_t)end1 * sizeof(double);
    // packed type 1 (la + lb <= kAsmPackLab): one slice of LDS per triplet
    {
      AsmArgs &a4 = P.aa4;
      const int LS = std::min(LM, kAsmPackLab), LQ = std::min(2 * LM, kAsmPackLab);
      a4.pair_chunk = aa.pair_chunk;
      a4.lds_rs_stride = (LQ + 1) * (LQ + 1);
      a4.shift_lm = LS;
      int o4 = 0;
      a4.lds_J = o4; o4 += ECP_DIM_IJK(LS) * ECP_DIM_C(LS);
      a4.lds_H = o4; o4 += ECP_DIM_C(LQ);
      a4.lds_S1 = o4;
      a4.lds_rsph = o4; o4 += a4.pair_chunk * a4.lds_rs_stride;
      a4.lds_Q = o4; o4 += a4.pair_chunk * nq_of(LQ);
      a4.lds_group = std::max(o4, a4.lds_S1 + shift_tab_doubles(LS));
      P.asm1_lds4 = (size_t)kAsmPack * a4.lds_group * sizeof(double);
    }
    // packed type 2 likewise: blk stays for the shift, whose tables reuse T
    {
      AsmArgs &a2 = P.aa2p;
      const int LS = std::min(LM, kAsmPackLab);
      a2.shift_lm = LS;
      int o2 = 0;
      a2.lds_J = o2; o2 += ECP_DIM_IJK(LS) * ECP_DIM_C(LS);
      a2.lds_blk = o2; o2 += ECP_DIM_C(LS) * ECP_DIM_C(LS);
      a2.lds_S2 = o2;
      a2.lds_T = o2; o2 += (LS + LC) * (LS + LC) * (2 * LS + 1);
      a2.lds_group = std::max(o2, a2.lds_S2 + shift_tab_doubles(LS));
      P.asm2_lds4 = (size_t)kAsmPack * a2.lds_group * sizeof(double);
    }
    P.asm2_lds = (size_t)end2 * sizeof(double);
    if (std::max(P.asm1_lds, P.asm2_lds) > (size_t)d->lds_max) {
      snprintf(err, errlen, "assembly needs %zu bytes of LDS, the device has %d per workgroup",
               std::max(P.asm1_lds, P.asm2_lds), d->lds_max);
      return -1;
    }
  }
  if (Ly.in_bytes > d->pin_in_cap) {
    if (d->pin_in) ECP_CHECK(hipHostFree(d->pin_in));
    d->pin_in = nullptr;
    d->pin_in_cap = 0;
    ECP_CHECK(hipHostMalloc(&d->pin_in, Ly.in_bytes + Ly.in_bytes / 4));
    d->pin_in_cap = Ly.in_bytes + Ly.in_bytes / 4;
  }
  const size_t out_bytes = 8 * job->nblk + 2 * (size_t)nt;
  if (out_bytes > d->pin_out_cap) {
    if (d->pin_out) ECP_CHECK(hipHostFree(d->pin_out));
    d->pin_out = nullptr;
    d->pin_out_cap = 0;
    ECP_CHECK(hipHostMalloc(&d->pin_out, out_bytes + out_bytes / 4));
    d->pin_out_cap = out_bytes + out_bytes / 4;
  }
  char *H = d->pin_in;
  auto cp = [&](size_t o, const void *src, size_t b) {
    if (b) memcpy(H + o, src, b);
  };
  cp(Ly.i_atoms, job->atoms, sizeof(ecp_xyz) * job->nat);
  cp(Ly.i_amaxl, job->atom_maxl, 4 * (size_t)job->nat);
  cp(Ly.i_sh, job->shells, sizeof(ecp_shell) * job->nsh);
  cp(Ly.i_pa, job->pa, 8 * (size_t)job->nprim);
  cp(Ly.i_pd, job->pd, 8 * (size_t)job->nprim);
  cp(Ly.i_cent, job->cent, sizeof(ecp_centre) * job->ncent);
  cp(Ly.i_gauss, job->gauss, sizeof(ecp_gauss) * job->ngauss);
  cp(Ly.i_win, job->win, sizeof(ecp_window) * job->nwin);
  cp(Ly.i_trip, job->trip, sizeof(ecp_triplet) * nt);
  cp(Ly.i_u1, units1.data(), sizeof(T1Unit) * units1.size());
  cp(Ly.i_t1f, t1first.data(), 4 * (size_t)nt);
  cp(Ly.i_grp, groups.data(), sizeof(T2Group) * groups.size());
  cp(Ly.i_goff, goff.data(), 4 * goff.size());
  cp(Ly.i_tgf, tgfirst.data(), 4 * (size_t)nt);
  cp(Ly.i_tcls2, tcls2.data(), (size_t)nt);
  cp(Ly.i_fit, fitems.data(), sizeof(int2) * fitems.size());
  cp(Ly.i_ait, aitems.data(), sizeof(int2) * aitems.size());
  cp(Ly.i_tcls, tcls.data(), 4 * (size_t)nt);
  if (bytes) *bytes = Ly.total;
  d->planned = true;
  return 0;
}

// Device half: places the planned job at `mem` (capacity cap; NULL = the
// context's own grow-only arena), uploads the staged inputs and waits.
extern "C" int ecp_dev_bind(ecp_dev *d, void *mem, size_t cap, char *err, size_t errlen) {
  if (!d->planned) {
    snprintf(err, errlen, "ecp_dev_bind before ecp_dev_plan");
    return -1;
  }
  Prepared &P = d->P;
  const JobLayout &Ly = d->L;
  d->prepared = false;
  d->eager_done = false;
  reset_graph(d);
  if (P.nt == 0) {
    d->prepared = true;
    return 0;
  }
  char *A = (char *)mem;
  if (!A) {
    if (Ly.total > d->arena_cap) {
      if (d->arena) ECP_CHECK(hipFree(d->arena));
      d->arena = nullptr;
      d->arena_cap = 0;
      const size_t c2 = Ly.total + Ly.total / 4;
      ECP_CHECK(hipMalloc(&d->arena, c2));
      d->arena_cap = c2;
    }
    A = d->arena;
  } else if (Ly.total > cap) {
    snprintf(err, errlen, "job needs %zu bytes of device memory, %zu given", Ly.total, cap);
    return -1;
  }
  ECP_CHECK(hipMemcpyAsync(A, d->pin_in, Ly.in_bytes, hipMemcpyHostToDevice, d->stream));
  ECP_CHECK(hipStreamSynchronize(d->stream));
  DevJob &J = P.J;
  J.atoms = (const ecp_xyz *)(A + Ly.i_atoms);
  J.atom_maxl = (const int *)(A + Ly.i_amaxl);
  J.shells = (const ecp_shell *)(A + Ly.i_sh);
  J.pa = (const double *)(A + Ly.i_pa);
  J.pd = (const double *)(A + Ly.i_pd);
  J.cent = (const ecp_centre *)(A + Ly.i_cent);
  J.gauss = (const ecp_gauss *)(A + Ly.i_gauss);
  J.win = (const ecp_window *)(A + Ly.i_win);
  J.trip = (const ecp_triplet *)(A + Ly.i_trip);
  J.ntrip = P.nt;
  J.UL = (double *)(A + Ly.d_UL);
  J.Ut = (double *)(A + Ly.d_Ut);
  J.F = (double *)(A + Ly.d_F);
  J.ang = (double *)(A + Ly.d_ang);
  J.FL = Ly.FL;
  J.UK = Ly.UK;
  J.angL = Ly.angL;
  J.ang_lam = Ly.ang_lam;
  J.ang_c = Ly.ang_c;
  J.deriv = Ly.deriv;
  J.maxPairs = P.maxPairs;
  P.u1 = (const T1Unit *)(A + Ly.i_u1);
  P.q = (double *)(A + Ly.d_q);
  P.t1s = (double *)(A + Ly.d_t1s);
  P.tq1 = (int *)(A + Ly.d_tq1);
  P.qf = (unsigned long long *)(A + Ly.d_qf);
  P.qfb = (double *)(A + Ly.d_qfb);
  P.qopen = (unsigned long long *)(A + Ly.d_qopen);
  P.qfbf = (unsigned long long *)(A + Ly.d_qfbf);
  P.udir = (double *)(A + Ly.d_udir);
  P.fbq1 = (int *)(A + Ly.d_fbq1);
  P.fbc = (int *)(A + Ly.d_fbc);  // [0..12] type-1 fallback queues, [32..44] type-1 tails, [48] type-2
                                  // tail, [kFb2Cnt + 0..12] type-2 fallback queues
  P.stat = (unsigned char *)(A + Ly.d_stat);
  P.t1fb = (unsigned char *)(A + Ly.d_t1fb);
  P.t2fb = (unsigned char *)(A + Ly.d_t2fb);
  P.grp = (const T2Group *)(A + Ly.i_grp);
  P.goff = (const int *)(A + Ly.i_goff);
  P.Tout = (double *)(A + Ly.d_T);
  P.tf = (unsigned char *)(A + Ly.d_tf);
  P.gflag = (int *)(A + Ly.d_gflag);
  P.flist = (int *)(A + Ly.d_flist);
  P.tq2 = (int *)(A + Ly.d_tq2);
  P.fbq2 = (int *)(A + Ly.d_fbq2);
  P.tclaim = (int *)(A + Ly.d_tclaim);
  P.fbf2 = (unsigned char *)(A + Ly.d_fbf2);
  P.tcls2 = (const unsigned char *)(A + Ly.i_tcls2);
  P.tgf = (const int *)(A + Ly.i_tgf);
  P.fbT = (double *)(A + Ly.d_fbT);
  P.fit = (const int2 *)(A + Ly.i_fit);
  P.ait = (const int2 *)(A + Ly.i_ait);
  P.tcls = (const int *)(A + Ly.i_tcls);
  P.blk = (double *)(A + Ly.d_blk);
  P.zero_base = (unsigned int *)(A + Ly.d_zero);
  P.zero_words = (Ly.d_zero_end - Ly.d_zero) / 4;
  for (AsmArgs *x : {&P.aa, &P.aa4, &P.aa2p}) {
    AsmArgs &aa = *x;
    aa.units1 = P.u1;
    aa.udir = P.udir;
    aa.t1first = (const int *)(A + Ly.i_t1f);
    aa.qout = P.q;
    aa.groups = P.grp;
    aa.tgfirst = (const int *)(A + Ly.i_tgf);
    aa.tmpl = d->tmpl_dev;
    aa.Tout = P.Tout;
    aa.tfail = P.tf;
    aa.fbT = P.fbT;
    aa.fac = d->fac_dev;
    aa.dfac = d->dfac_dev;
    aa.blocks = P.blk;
  }
  d->prepared = true;
  return 0;
}

extern "C" int ecp_dev_prepare(ecp_dev *d, const ecp_job *job, char *err, size_t errlen) {
  if (ecp_dev_plan(d, job, nullptr, err, errlen)) return -1;
  return ecp_dev_bind(d, nullptr, 0, err, errlen);
}

// Runs the prepared evaluation `repeats` times and waits.  The first
// evaluation of a job runs eagerly and records the stage timings; when more
// evaluations follow (in this call or a later one) the kernel sequence is
// captured once into a HIP graph and replayed.
extern "C" int ecp_dev_execute(ecp_dev *d, int repeats, double *ms_total, char *err, size_t errlen) {
  if (!d->prepared) {
    snprintf(err, errlen, "ecp_dev_execute before ecp_dev_prepare");
    return -1;
  }
  d->nev = 0;
  if (ms_total) *ms_total = 0.0;
  if (d->P.nt == 0 || repeats <= 0) return 0;
  hipStream_t s = d->stream;
  int remaining = repeats;
  double total = 0.0;
  d->last_eager = false;
  if (!d->eager_done) {
    // a stale error of an earlier, unrelated call must not be taken for one
    // of this enqueue's launches
    (void)hipGetLastError();
    d->span_k = 0;
    d->last_eager = true;
    enqueue(d, s, d->ev);
    ECP_CHECK(hipGetLastError());
    ECP_CHECK(hipStreamSynchronize(s));
    for (int i = 1; i < 4; i++) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, d->ev[i - 1], d->ev[i]);
      d->last_ms[i - 1] = ms;
    }
    d->nev = 3;
    total = d->last_ms[0] + d->last_ms[1] + d->last_ms[2];
    d->eager_done = true;
    remaining--;
  }
  if (remaining > 0 && d->ngraph == 0) {
    // fused launches: kSpanGraphs copies, each recording its own quadrature
    // event pair; per-class launches: one graph
    for (int k = 0; k < kSpanGraphs; k++) {
      d->span_k = k;
      // a first host-side record, so that the pair reads as recorded before
      // the graph's event nodes overwrite it
      ECP_CHECK(hipEventRecord(d->span[k][0], s));
      ECP_CHECK(hipEventRecord(d->span[k][1], s));
      ECP_CHECK(hipStreamSynchronize(s));
      ECP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      d->capturing = true;
      enqueue(d, s, nullptr);
      d->capturing = false;
      ECP_CHECK(hipStreamEndCapture(s, &d->graph[k]));
      ECP_CHECK(hipGraphInstantiate(&d->gexec[k], d->graph[k], nullptr, nullptr, 0));
      d->ngraph = k + 1;
      if (!d->span_valid) break;
    }
    d->span_k = 0;
    d->replays = 0;
  }
  if (remaining > 0) {
    d->last_eager = false;
    d->last_replays = remaining;
    ECP_CHECK(hipEventRecord(d->ev[5], s));
    for (int r = 0; r < remaining; r++) {
      ECP_CHECK(hipGraphLaunch(d->gexec[d->replays % d->ngraph], s));
      d->replays++;
    }
    ECP_CHECK(hipEventRecord(d->ev[6], s));
    ECP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, d->ev[5], d->ev[6]);
    total = ms / remaining;
    d->last_ms[3] = (float)total;
    d->nev = 4;
  }
  if (ms_total) *ms_total = total;
  return 0;
}

// One eager evaluation with an event pair around every kernel launch on its
// own stream; per family: launches, summed duration (ms), algorithmic FLOPs.
namespace {
// Algorithmic FLOPs of the large-grid fallbacks of the last evaluation
// (ecp_fb1_flops / ecp_fb2_flops per failed primitive pair / group): every
// term is an integer below 2^53, so the atomic sum is exact in any order.
__global__ __launch_bounds__(256) void k_fb_flops(DevJob J, const T1Unit *__restrict__ units,
                                                  const unsigned long long *__restrict__ qf, int nunits,
                                                  const T2Group *__restrict__ groups, const int *__restrict__ gflag,
                                                  const int *__restrict__ flist, const unsigned char *__restrict__ tfail,
                                                  int ngroups, int N, double *out) {
  double s1 = 0.0, s2 = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < nunits; u += stride) {
    const int f = __popcll(qf[u]);
    if (!f) continue;
    const ecp_triplet tr = J.trip[units[u].trip];
    s1 += ecp_fb1_flops(tr.la + tr.lb, f, N);
  }
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    // the group's failed entries among those listed for the fallback (a
    // speculative fallback also lists entries the small grid then decided)
    const T2Group G = groups[g];
    int f = 0;
    for (int k = 0; k < gflag[g]; k++) f += tfail[G.eoff + flist[G.eoff + k]];
    if (!f) continue;
    const ecp_triplet tr = J.trip[G.trip];
    const ecp_centre C = J.cent[tr.centre];
    int ng = 0;
    for (int i = 0; i < C.ng; i++) ng += J.gauss[C.g0 + i].l == G.l;
    const int KK = J.shells[tr.sa].K * J.shells[tr.sb].K;
    s2 += KK * ecp_fb2_flops(tr.la, tr.lb, G.l, ng, f, N);
  }
  if (s1 != 0.0) atomicAdd(&out[0], s1);
  if (s2 != 0.0) atomicAdd(&out[1], s2);
}

}  // namespace

extern "C" int ecp_dev_profile(ecp_dev *d, int *launches, double *ms, double *flops, char *err, size_t errlen) {
  if (!d->prepared) {
    snprintf(err, errlen, "ecp_dev_profile before ecp_dev_prepare");
    return -1;
  }
  for (int f = 0; f < FAM_N; f++) {
    launches[f] = 0;
    ms[f] = 0.0;
    flops[f] = 0.0;
  }
  if (d->P.nt == 0) return FAM_N;
  d->profiling = true;
  d->npev = 0;
  d->span_k = 0;
  d->last_eager = true;
  (void)hipGetLastError();
  enqueue(d, d->stream, nullptr);
  d->profiling = false;
  ECP_CHECK(hipGetLastError());
  ECP_CHECK(hipStreamSynchronize(d->stream));
  for (int i = 0; i < d->npev; i++) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, d->pev[2 * i], d->pev[2 * i + 1]);
    launches[d->pfam[i]]++;
    ms[d->pfam[i]] += t;
  }
  // algorithmic FLOPs per family: the model's terms known before the
  // evaluation, and the fallbacks from this evaluation's failed integrals
  const Prepared &P = d->P;
  flops[FAM_TABLES] = P.model[ECP_FL_FTAB];
  flops[FAM_TYPE1] = P.model[ECP_FL_T1];
  flops[FAM_TYPE2] = P.model[ECP_FL_T2];
  flops[FAM_ASM] = P.model[ECP_FL_ASM1] + P.model[ECP_FL_ASM2] + P.model[ECP_FL_SHIFT];
  if (!d->fbf) ECP_CHECK(hipMalloc(&d->fbf, 2 * sizeof(double)));
  ECP_CHECK(hipMemsetAsync(d->fbf, 0, 2 * sizeof(double), d->stream));
  const int nunits = P.cls_begin[ECP_MAX_LAB + 1];
  const int nb = std::max(1, std::min(1024, (std::max(nunits, P.ngroups) + 255) / 256));
  k_fb_flops<<<nb, 256, 0, d->stream>>>(P.J, P.u1, P.qf, nunits, P.grp, P.gflag, P.flist, P.tf, P.ngroups,
                                        d->T.lorder, d->fbf);
  ECP_CHECK(hipGetLastError());
  double fb[2] = {0.0, 0.0};
  ECP_CHECK(hipMemcpyAsync(fb, d->fbf, sizeof(fb), hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipStreamSynchronize(d->stream));
  flops[FAM_FB1] = fb[0];
  flops[FAM_FB2] = fb[1];
  return FAM_N;
}

extern "C" const char *ecp_dev_family_name(int f) { return (f >= 0 && f < FAM_N) ? kFamName[f] : ""; }

extern "C" int ecp_dev_fetch(ecp_dev *d, double *blocks_out, unsigned char *status_out, char *err, size_t errlen) {
  const Prepared &P = d->P;
  if (P.nt == 0) return 0;
  ECP_CHECK(hipMemcpyAsync(d->pin_out, P.blk, 8 * P.nblk, hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipMemcpyAsync(d->pin_out + 8 * P.nblk, P.stat, 2 * (size_t)P.nt, hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipStreamSynchronize(d->stream));
  memcpy(blocks_out, d->pin_out, 8 * P.nblk);
  memcpy(status_out, d->pin_out + 8 * P.nblk, 2 * (size_t)P.nt);
  return 0;
}

extern "C" int ecp_dev_run(ecp_dev *d, const ecp_job *job, double *blocks_out, unsigned char *status_out,
                           char *err, size_t errlen) {
  if (ecp_dev_prepare(d, job, err, errlen)) return -1;
  if (ecp_dev_execute(d, 1, nullptr, err, errlen)) return -1;
  return ecp_dev_fetch(d, blocks_out, status_out, err, errlen);
}

// ---------------------------------------------------------------------------
// Density contraction of the evaluated blocks on the device, the consumer
// that getIntegrals.c:15-18 anticipates (a density P and a gradient in
// place of the AO matrix): no blocks leave the device.
//   n = 0: E_m = sum_{mu nu} P_{mu nu} V_{mu nu}; a block of shells a != b
//          stands for itself and its transpose (weight 2), a diagonal shell
//          block (a = b) for itself (weight 1).
//   n = 1: dE_m/dR_{X,d}.  A primitive's derivative is
//          d/dA_d g(a) = 2 zeta g(a + 1_d) - a_d g(a - 1_d) (libecp.c:242-245),
//          whose zeta factor the (+1) blocks already carry; the potential's
//          own derivative follows from translational invariance,
//          d/dC = -(d/dA + d/dB), so each block's contribution s goes to the
//          atom of its shifted shell with +s and to the ECP centre with -s.
//          Blocks the screening skips (A = B = C; the side on the centre,
//          libecp.c:305-306,326-330) would cancel in that sum.
// Partial sums per triplet are reduced by a fixed butterfly, and the
// per-molecule / per-atom sums run over triplets in a fixed order, so
// results do not depend on scheduling.

namespace {
struct ConArgs {
  const ecp_triplet *trip;
  const ecp_shell *shells;
  const ecp_centre *cent;
  const double *blocks;
  const unsigned char *stat;
  const int *aoff;        // AO offset of each global shell within its molecule
  const double *P;        // densities, packed by molecule
  const long long *poff;  // [nmol] offset of each molecule's density
  const int *prow;        // [nmol] its row stride
  const int *tmb;         // [nmol + 1] triplet range of each molecule
  const int *umol;        // [nunits] molecule of each output unit
  const int *uatom;       // [nunits] global atom of each output unit (n = 1)
  const int *uout;        // [nunits] output offset of each unit
  double *part;           // [ntrip * nd] weighted partial sums
  double *out;
  int *codes;             // [nmol] first failed integration: 0, 1 or 2
  int ntrip, nd, nunits, nmol;
};

__device__ __forceinline__ double wave_sum_fixed(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(64) void k_con_part(DevTables T, ConArgs A) {
  const int lane = threadIdx.x;
  const int D = T.ijkDim;
  for (int t = blockIdx.x; t < A.ntrip; t += gridDim.x) {
    const ecp_triplet tr = A.trip[t];
    const ecp_shell Sa = A.shells[tr.sa], Sb = A.shells[tr.sb];
    const double *__restrict__ Pm = A.P + A.poff[Sa.mol];
    const int R = A.prow[Sa.mol];
    const int a0 = A.aoff[tr.sa], b0 = A.aoff[tr.sb];
    const int la = Sa.l, lb = Sb.l, nb = ECP_DIM_IJK(lb), nab = ECP_DIM_IJK(la) * nb;
    const int no = ECP_DIM_IJK(tr.ob);
    const double *__restrict__ b1 = A.blocks + tr.blk1;
    const double *__restrict__ b2 = A.blocks + tr.blk2;
    const bool bra = tr.shifta != 0;
    const int sh = bra ? tr.shifta : tr.shiftb, ls = bra ? la : lb;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = lane; e < nab; e += 64) {
      const int i = e / nb, j = e - i * nb;
      const double p = Pm[(size_t)(a0 + i) * R + b0 + j];
      if (A.nd == 1) {
        acc[0] += p * (b1[i * no + j] + b2[i * no + j]);
        continue;
      }
      const int *ex = T.ijk + 3 * ECP_IDX_C(ls, bra ? i : j);
#pragma unroll
      for (int dd = 0; dd < 3; dd++) {
        int x[3] = {ex[0], ex[1], ex[2]};
        double coef;
        if (sh > 0) {
          coef = 2.0;
          x[dd] += 1;
        } else {
          if (x[dd] == 0) continue;
          coef = -(double)x[dd];
          x[dd] -= 1;
        }
        const int c = T.ijkIndex[(x[0] * D + x[1]) * D + x[2]] - ECP_IDX_C(ls + sh, 0);
        const int o = bra ? c * no + j : i * no + c;
        acc[dd] += coef * (p * (b1[o] + b2[o]));
      }
    }
    const double w = (tr.sa == tr.sb) ? 1.0 : 2.0;
    for (int k = 0; k < A.nd; k++) {
      const double v = wave_sum_fixed(acc[k]);
      if (lane == 0) A.part[(size_t)t * A.nd + k] = w * v;
    }
  }
}

constexpr int kConThreads = 256;

__global__ __launch_bounds__(kConThreads) void k_con_reduce(ConArgs A) {
  __shared__ double red[3][kConThreads];
  __shared__ int fmin[kConThreads];
  const int u = blockIdx.x, tid = threadIdx.x;
  const int m = A.umol[u], g = A.uatom[u];
  const int t0 = A.tmb[m], t1 = A.tmb[m + 1];
  double s[3] = {0.0, 0.0, 0.0};
  int first = INT_MAX;
  for (int t = t0 + tid; t < t1; t += kConThreads) {
    if (A.stat[2 * t]) first = min(first, 2 * t);
    else if (A.stat[2 * t + 1]) first = min(first, 2 * t + 1);
    if (A.nd == 1) {
      s[0] += A.part[t];
      continue;
    }
    const ecp_triplet tr = A.trip[t];
    const int X = A.shells[tr.shifta != 0 ? tr.sa : tr.sb].atom, C = A.cent[tr.centre].atom;
    const double *v = A.part + (size_t)t * 3;
    if (X == g)
      for (int k = 0; k < 3; k++) s[k] += v[k];
    if (C == g)
      for (int k = 0; k < 3; k++) s[k] -= v[k];
  }
  for (int k = 0; k < 3; k++) red[k][tid] = s[k];
  fmin[tid] = first;
  __syncthreads();
  for (int w = kConThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      for (int k = 0; k < 3; k++) red[k][tid] += red[k][tid + w];
      fmin[tid] = min(fmin[tid], fmin[tid + w]);
    }
    __syncthreads();
  }
  if (tid < A.nd) A.out[A.uout[u] + tid] = red[tid][0];
  // one unit per molecule reports its first failure (callback order)
  if (tid == 0 && (u == 0 || A.umol[u - 1] != m))
    A.codes[m] = fmin[0] == INT_MAX ? 0 : 1 + (fmin[0] & 1);
}
}  // namespace

namespace {

// Index tables of a contraction (device layout, one block of `bytes`): the
// density offsets and row strides per molecule, the AO offset of every
// shell, the triplet range of every molecule, and the output units (one
// per molecule for the energy, one per atom for the gradient: molecule,
// global atom, output offset).  img == nullptr: sizes only.
struct ConImage {
  size_t poff = 0, aoff = 0, prow = 0, tmb = 0, umol = 0, uatom = 0, uout = 0, bytes = 0;
  int nunits = 0;
  size_t nout = 0;
};

ConImage con_image(const ecp_job *job, const long long *poff, const int *prow, char *img) {
  ConImage c;
  const int nmol = job->nmol, nd = job->deriv == 0 ? 1 : 3;
  c.nunits = nd == 1 ? nmol : job->mol_atom0[nmol];
  c.nout = (size_t)c.nunits * nd;
  size_t q = 0;
  auto take = [&](size_t bytes) {
    const size_t o = q;
    q += (bytes + 7) & ~(size_t)7;
    return o;
  };
  c.poff = take(8 * (size_t)nmol);
  c.aoff = take(4 * ((size_t)job->nsh + 1));
  c.prow = take(4 * (size_t)nmol);
  c.tmb = take(4 * ((size_t)nmol + 1));
  c.umol = take(4 * (size_t)c.nunits);
  c.uatom = take(4 * (size_t)c.nunits);
  c.uout = take(4 * (size_t)c.nunits);
  c.bytes = q;
  if (!img) return c;
  memcpy(img + c.poff, poff, 8 * (size_t)nmol);
  ecp_job_ao_offsets(job, (int *)(img + c.aoff));
  memcpy(img + c.prow, prow, 4 * (size_t)nmol);
  memcpy(img + c.tmb, job->trip_mol_begin, 4 * ((size_t)nmol + 1));
  int *um = (int *)(img + c.umol), *ua = (int *)(img + c.uatom), *uo = (int *)(img + c.uout);
  int u = 0;
  for (int m = 0; m < nmol; m++) {
    if (nd == 1) {
      um[u] = m;
      ua[u] = -1;
      uo[u] = u;
      u++;
      continue;
    }
    for (int g = job->mol_atom0[m]; g < job->mol_atom0[m + 1]; g++, u++) {
      um[u] = m;
      ua[u] = g;
      uo[u] = 3 * u;
    }
  }
  return c;
}

// Launches the contraction over a prepared image at base + o_idx, partial
// sums at part, codes at codes_dev; P and out are device pointers.
int con_launch(ecp_dev *d, const ecp_job *job, const ConImage &c, char *idx, const double *P, double *part,
               double *out, int *codes_dev, char *err, size_t errlen) {
  const Prepared &Pp = d->P;
  ConArgs A{};
  A.trip = Pp.J.trip;
  A.shells = Pp.J.shells;
  A.cent = Pp.J.cent;
  A.blocks = Pp.blk;
  A.stat = Pp.stat;
  A.aoff = (const int *)(idx + c.aoff);
  A.P = P;
  A.poff = (const long long *)(idx + c.poff);
  A.prow = (const int *)(idx + c.prow);
  A.tmb = (const int *)(idx + c.tmb);
  A.umol = (const int *)(idx + c.umol);
  A.uatom = (const int *)(idx + c.uatom);
  A.uout = (const int *)(idx + c.uout);
  A.part = part;
  A.out = out;
  A.codes = codes_dev;
  A.ntrip = Pp.nt;
  A.nd = job->deriv == 0 ? 1 : 3;
  A.nunits = c.nunits;
  A.nmol = job->nmol;
  ECP_CHECK(hipEventRecord(d->con_ev[0], d->stream));
  if (Pp.nt > 0) {
    k_con_part<<<std::min(Pp.nt, 65536), 64, 0, d->stream>>>(d->T, A);
    ECP_CHECK(hipGetLastError());
  }
  if (c.nunits > 0) {
    k_con_reduce<<<c.nunits, kConThreads, 0, d->stream>>>(A);
    ECP_CHECK(hipGetLastError());
  }
  ECP_CHECK(hipEventRecord(d->con_ev[1], d->stream));
  return 0;
}

}  // namespace

extern "C" int ecp_dev_contract(ecp_dev *d, const ecp_job *job, const double *P, size_t nP, int P_on_device,
                                const long long *poff, const int *prow, double *out, int out_on_device,
                                int *codes, double *ms, char *err, size_t errlen) {
  const Prepared &Pp = d->P;
  if (!d->prepared) {
    snprintf(err, errlen, "no evaluated job to contract");
    return -1;
  }
  if (job->deriv > 1) {
    snprintf(err, errlen, "density contraction supports derivative orders 0 and 1");
    return -1;
  }
  const int nmol = job->nmol, nd = job->deriv == 0 ? 1 : 3;
  const ConImage ci = con_image(job, poff, prow, nullptr);
  std::vector<char> h(ci.bytes);
  con_image(job, poff, prow, h.data());
  // device layout
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  const size_t o_part = o; o += al(8 * (size_t)std::max(1, Pp.nt) * nd);
  const size_t o_out = o; o += al(8 * std::max<size_t>(1, ci.nout));
  const size_t o_codes = o; o += al(4 * (size_t)nmol);
  const size_t o_idx = o; o += al(ci.bytes);
  const size_t o_P = o; o += P_on_device ? 0 : al(8 * nP);
  if (o > d->con_cap) {
    if (d->con_mem) ECP_CHECK(hipFree(d->con_mem));
    d->con_mem = nullptr;
    d->con_cap = 0;
    ECP_CHECK(hipMalloc(&d->con_mem, o));
    d->con_cap = o;
    d->con_idx.clear();
  }
  char *b = d->con_mem;
  // the index tables change with the job or the density layout only
  if (d->con_idx_at != o_idx || d->con_idx != h) {
    ECP_CHECK(hipMemcpyAsync(b + o_idx, h.data(), h.size(), hipMemcpyHostToDevice, d->stream));
    ECP_CHECK(hipStreamSynchronize(d->stream));
    d->con_idx = h;
    d->con_idx_at = o_idx;
  }
  if (!P_on_device && nP) ECP_CHECK(hipMemcpyAsync(b + o_P, P, 8 * nP, hipMemcpyHostToDevice, d->stream));
  double *outp = out_on_device ? out : (double *)(b + o_out);
  if (con_launch(d, job, ci, b + o_idx, P_on_device ? P : (const double *)(b + o_P), (double *)(b + o_part), outp,
                 (int *)(b + o_codes), err, errlen))
    return -1;
  if (!out_on_device && ci.nout)
    ECP_CHECK(hipMemcpyAsync(out, b + o_out, 8 * ci.nout, hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipMemcpyAsync(codes, b + o_codes, 4 * (size_t)nmol, hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipStreamSynchronize(d->stream));
  if (ms) {
    float f = 0.f;
    ECP_CHECK(hipEventElapsedTime(&f, d->con_ev[0], d->con_ev[1]));
    *ms = f;
  }
  return 0;
}

extern "C" size_t ecp_dev_contract_image(const ecp_job *job, const long long *poff, const int *prow, char *img,
                                         size_t cap) {
  const ConImage c = con_image(job, poff, prow, nullptr);
  if (img && cap >= c.bytes) con_image(job, poff, prow, img);
  return c.bytes;
}

extern "C" int ecp_dev_contract_noalloc(ecp_dev *d, const ecp_job *job, const char *img, size_t img_bytes,
                                        const double *P, double *out, int *codes, double *ms, char *err,
                                        size_t errlen) {
  const Prepared &Pp = d->P;
  if (!d->prepared) {
    snprintf(err, errlen, "no evaluated job to contract");
    return -1;
  }
  if (job->deriv > 1) {
    snprintf(err, errlen, "density contraction supports derivative orders 0 and 1");
    return -1;
  }
  if (!d->scratch) {
    snprintf(err, errlen, "the no-allocation scratch of this device is not set up");
    return -1;
  }
  const int nmol = job->nmol, nd = job->deriv == 0 ? 1 : 3;
  // the image's sizes, from the same layout function that wrote it (no
  // allocation: sizes only)
  const ConImage ci = con_image(job, nullptr, nullptr, nullptr);
  if (ci.bytes != img_bytes) {
    snprintf(err, errlen, "contraction index image of %zu bytes does not match the job (%zu)", img_bytes, ci.bytes);
    return -1;
  }
  // scratch layout: partial sums, codes, index tables
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  const size_t o_part = o; o += al(8 * (size_t)std::max(1, Pp.nt) * nd);
  const size_t o_codes = o; o += al(4 * (size_t)nmol);
  const size_t o_idx = o; o += al(ci.bytes);
  if (o > kScratchBytes) {
    snprintf(err, errlen,
             "the contraction needs %zu bytes of scratch (%d triplets, %d molecules); the no-allocation "
             "entry point holds %zu",
             o, Pp.nt, nmol, kScratchBytes);
    return -1;
  }
  char *b = (char *)d->scratch;
  // nothing in the scratch has to start at zero: the kernels write every
  // partial sum, output and code they read
  ECP_CHECK(hipMemcpyAsync(b + o_idx, img, ci.bytes, hipMemcpyHostToDevice, d->stream));
  if (con_launch(d, job, ci, b + o_idx, P, (double *)(b + o_part), out, (int *)(b + o_codes), err, errlen))
    return -1;
  if (codes) ECP_CHECK(hipMemcpyAsync(codes, b + o_codes, 4 * (size_t)nmol, hipMemcpyDeviceToHost, d->stream));
  ECP_CHECK(hipStreamSynchronize(d->stream));
  if (ms) {
    float f = 0.f;
    ECP_CHECK(hipEventElapsedTime(&f, d->con_ev[0], d->con_ev[1]));
    *ms = f;
  }
  return 0;
}

extern "C" int ecp_dev_span_ms(ecp_dev *d, double *ms, int *samples, char *err, size_t errlen) {
  if (!d->prepared || !d->span_valid || (!d->last_eager && d->replays == 0)) {
    snprintf(err, errlen, "no single-stream type-1 quadrature span in the last evaluation");
    return -1;
  }
  // the eager evaluation, or the last replays (one per graph copy)
  const int n = d->last_eager ? 1 : std::min(d->ngraph, d->last_replays);
  double sum = 0.0;
  for (int i = 0; i < n; i++) {
    const int k = d->last_eager ? 0 : (int)((d->replays - 1 - i) % d->ngraph);
    float f = 0.f;
    ECP_CHECK(hipEventElapsedTime(&f, d->span[k][0], d->span[k][1]));
    sum += f;
  }
  *ms = sum / n;
  if (samples) *samples = n;
  return 0;
}

extern "C" int ecp_dev_mem_alloc(void **p, size_t bytes, char *err, size_t errlen) {
  *p = nullptr;
  ECP_CHECK(hipMalloc(p, bytes));
  return 0;
}

extern "C" void ecp_dev_mem_free(void *p) {
  if (p) (void)hipFree(p);
}

extern "C" int ecp_dev_sync(ecp_dev *d) {
  int rc = 0;
  if (d->stream && hipStreamSynchronize(d->stream) != hipSuccess) rc = -1;
  for (int k = 0; k < kWorkers; k++)
    if (d->workers[k] && hipStreamSynchronize(d->workers[k]) != hipSuccess) rc = -1;
  return rc;
}

extern "C" int ecp_dev_current(void) {
  int dv = 0;
  return hipGetDevice(&dv) == hipSuccess ? dv : -1;
}

extern "C" int ecp_dev_use(int device) { return hipSetDevice(device) == hipSuccess ? 0 : -1; }

extern "C" void ecp_dev_get_tuning(const ecp_dev *d, ecp_tuning *t) { *t = d->tune; }

extern "C" void ecp_dev_set_tuning(ecp_dev *d, ecp_tuning const *t) {
  d->tune = *t;
  d->tune.t1_tail_grid = std::max(1, d->tune.t1_tail_grid);
  reset_graph(d);
  d->eager_done = false;
}
