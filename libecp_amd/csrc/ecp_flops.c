/* Algorithmic FLOP model of the ECP-integral path (host side, C11).
 *
 * Counts the floating-point work of the reference's eager algorithm on a
 * screened job: every window point tabulated (type1.c:121-130,
 * type2.c:282-299), every adaptive integral over all its levels, every
 * large-grid fallback over all its points (type1.c:148-196,
 * type2.c:417-528), the Cartesian assemblies (type1.c:266-295,
 * type2.c:583-623) and the binomial shift (util.c:246-334).  Convention:
 * FMA = 2, add / sub / mul / div = 1, exp / pow / sqrt / log / trig = 1,
 * loads 0.  The kernels evaluate lazily (points only up to convergence), so
 * the model is reference-equivalent work, not executed instructions.
 *
 * Per-unit terms, G = window length ge - gs, lambda = la + lb,
 * n_Q(lambda) = sum_{l1 <= lambda} (floor(l1 / 2) + 1):
 *   F table, per (centre, shell s, derivative run), m = L - 1 + l_s + run:
 *       K_s G_s (30 m + 72 + 3 (m + 1))          Bessel 30 m + 68, exp and
 *                                                offset 4, F += d K e 3/order
 *   type 1 small grid, per primitive pair:
 *       G (31 lambda + 74 + 4 n_Q) + 150 n_Q
 *   type 1 assembly, per primitive pair:
 *       6 (lambda + 1)^2 + 36 + sum_{a in M(la), b in M(lb)}
 *                                sum_{l = deg(ab); l >= 0; l -= 2} (4 l + 4)
 *   type 2 small grid, per (triplet, l < L): |S_l| (4 G + 150)
 *   type 2 assembly, per (triplet, l):
 *       sum_{alpha <= la, beta <= lb} IJK(alpha) IJK(beta) n(alpha) n(beta) (4 l + 4)
 *       n(alpha) = number of lambda in [max((alpha + l) mod 2, l - alpha), l + alpha]
 *   shift, per triplet (both types):
 *       2 [sum_{|a| = la} prod(a_i + 1) (3 + 2 C_DIM(lb))
 *          + sum_{|b| = lb} prod(b_i + 1) (4 + 2 IJK(la))]
 *   type 1 fallback, per primitive pair with f failed entries, N = large-grid
 *   points: N (31 lambda + 74) + f (4 N + 150)
 *   type 2 fallback, per (triplet, l) with f failed entries, per primitive
 *   pair: N (30 (laC + lbC) + 136 + 8 + 5 g_l + lambda) + f (8 N + 150),
 *   laC = la + l, lbC = lb + l, g_l Gaussians of channel l: two Bessel
 *   vectors, the exponent and its exp, U_l, the powers r^k per point, and
 *   per entry and point the five-factor integrand, weight and sum.
 * M(l): the monomials of degree <= l; S_l: the (lambda1, lambda2, alpha +
 * beta) with lambda1 in n(alpha)'s range, lambda2 likewise with beta.
 */
#include "ecp_host.h"

static int nq_of(int lab) {
  int n = 0;
  for (int l1 = 0; l1 <= lab; l1++) n += l1 / 2 + 1;
  return n;
}

static double lam_count(int alpha, int l) {
  int lo = l - alpha, par = (alpha + l) % 2;
  if (par > lo) lo = par;
  const int hi = l + alpha;
  return hi >= lo ? (double)((hi - lo) / 2 + 1) : 0.0;
}

/* |S_l|: distinct (lambda1, lambda2, alpha + beta) over alpha <= la, beta <= lb */
static int s_size(int la, int lb, int l) {
  const int d1 = la + l + 1, d2 = lb + l + 1, d3 = la + lb + 1;
  unsigned char used[(ECP_MAX_LBS + ECP_MAX_L + 3) * (ECP_MAX_LBS + ECP_MAX_L + 3) * (2 * ECP_MAX_LBS + 5)];
  if ((size_t)d1 * d2 * d3 > sizeof(used)) return 0;
  for (int i = 0; i < d1 * d2 * d3; i++) used[i] = 0;
  int n = 0;
  for (int alpha = 0; alpha <= la; alpha++)
    for (int beta = 0; beta <= lb; beta++) {
      int lo1 = l - alpha, lo2 = l - beta;
      if ((alpha + l) % 2 > lo1) lo1 = (alpha + l) % 2;
      if ((beta + l) % 2 > lo2) lo2 = (beta + l) % 2;
      for (int l1 = lo1; l1 <= l + alpha; l1 += 2)
        for (int l2 = lo2; l2 <= l + beta; l2 += 2) {
          const int k = (l1 * d2 + l2) * d3 + alpha + beta;
          if (!used[k]) {
            used[k] = 1;
            n++;
          }
        }
    }
  return n;
}

double ecp_model_t1_unit(int lab, int G) {
  const int nq = nq_of(lab);
  return (double)G * (31.0 * lab + 74 + 4.0 * nq) + 150.0 * nq;
}

double ecp_model_asm1_unit(int la, int lb) {
  const int lab = la + lb;
  double s = 6.0 * (lab + 1) * (lab + 1) + 36;
  for (int da = 0; da <= la; da++)
    for (int db = 0; db <= lb; db++) {
      double t = 0.0;
      for (int l = da + db; l >= 0; l -= 2) t += 4.0 * l + 4;
      s += (double)ECP_DIM_IJK(da) * ECP_DIM_IJK(db) * t;
    }
  return s;
}

double ecp_model_t2_group(int la, int lb, int l, int G) { return (double)s_size(la, lb, l) * (4.0 * G + 150); }

double ecp_model_asm2_group(int la, int lb, int l) {
  double s = 0.0;
  for (int alpha = 0; alpha <= la; alpha++)
    for (int beta = 0; beta <= lb; beta++)
      s += (double)ECP_DIM_IJK(alpha) * ECP_DIM_IJK(beta) * lam_count(alpha, l) * lam_count(beta, l) *
           (4.0 * l + 4);
  return s;
}

double ecp_model_shift(int la, int lb) {
  double sa = 0.0, sb = 0.0;
  for (int ax = 0; ax <= la; ax++)
    for (int ay = 0; ax + ay <= la; ay++) sa += (double)(ax + 1) * (ay + 1) * (la - ax - ay + 1);
  for (int bx = 0; bx <= lb; bx++)
    for (int by = 0; bx + by <= lb; by++) sb += (double)(bx + 1) * (by + 1) * (lb - bx - by + 1);
  return 2.0 * (sa * (3 + 2.0 * ECP_DIM_C(lb)) + sb * (4 + 2.0 * ECP_DIM_IJK(la)));
}

double ecp_model_ftab(int K, int G, int m) { return (double)K * G * (30.0 * m + 72 + 3.0 * (m + 1)); }

double ecp_model_fb1_unit(int lab, int f, int N) { return ecp_fb1_flops(lab, f, N); }

double ecp_model_fb2_pair(int la, int lb, int l, int ng_l, int f, int N) {
  return ecp_fb2_flops(la, lb, l, ng_l, f, N);
}

void ecp_model_flops(const ecp_job *J, double *out) {
  for (int i = 0; i < ECP_FL_N; i++) out[i] = 0.0;
  /* F tables */
  for (int c = 0; c < J->ncent; c++) {
    const ecp_centre *C = &J->cent[c];
    for (int s = 0; s < C->nsh; s++) {
      const ecp_window *W = &J->win[C->win_off + s];
      if (W->skip) continue;
      const ecp_shell *S = &J->shells[C->shell0 + s];
      for (int run = 0; run <= J->deriv; run++)
        out[ECP_FL_FTAB] += ecp_model_ftab(S->K, W->end - W->start, C->L - 1 + S->l + run);
    }
  }
  /* the per-(la, lb[, l]) terms, formed once per job */
  enum { NL = ECP_MAX_LBS + 3 };
  double a1[NL][NL], sh[NL][NL], a2[NL][NL][ECP_MAX_L + 1];
  int ns[NL][NL][ECP_MAX_L + 1];
  for (int la = 0; la < NL; la++)
    for (int lb = 0; lb < NL; lb++) {
      a1[la][lb] = ecp_model_asm1_unit(la, lb);
      sh[la][lb] = ecp_model_shift(la, lb);
      for (int l = 0; l <= ECP_MAX_L; l++) {
        a2[la][lb][l] = ecp_model_asm2_group(la, lb, l);
        ns[la][lb][l] = s_size(la, lb, l);
      }
    }
  for (int t = 0; t < J->ntrip; t++) {
    const ecp_triplet *tr = &J->trip[t];
    if (tr->la >= NL || tr->lb >= NL || tr->oa >= NL || tr->ob >= NL) continue;
    const int KK = J->shells[tr->sa].K * J->shells[tr->sb].K;
    const int G = tr->ge - tr->gs;
    out[ECP_FL_T1] += KK * ecp_model_t1_unit(tr->la + tr->lb, G);
    out[ECP_FL_ASM1] += KK * a1[tr->la][tr->lb];
    for (int l = 0; l < J->cent[tr->centre].L && l <= ECP_MAX_L; l++) {
      out[ECP_FL_T2] += ns[tr->la][tr->lb][l] * (4.0 * G + 150);
      out[ECP_FL_ASM2] += a2[tr->la][tr->lb][l];
    }
    out[ECP_FL_SHIFT] += sh[tr->oa][tr->ob];
  }
}

int ecp_model_gauss_of(const ecp_job *J, int centre, int l) {
  const ecp_centre *C = &J->cent[centre];
  int n = 0;
  for (int g = 0; g < C->ng; g++) n += J->gauss[C->g0 + g].l == l;
  return n;
}
