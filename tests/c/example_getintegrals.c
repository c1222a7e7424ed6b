/* A C caller of the one-call interface, linked against libecp_amd.so the way
 * the reference's example links libecp.a (example/ex1.c:9,162): reads one
 * molecule in a whitespace-separated text format from stdin, calls
 * getIntegrals() and prints the upper triangle of the ECP matrix.
 *
 * Input: nat; geometry[3 nat]; shellsECP[nat]; nshECP; KECP[nshECP];
 * lECP[nshECP]; nprimECP; nECP, dECP, aECP [nprimECP each]; shellsBS[nat];
 * nsh; lBS[nsh]; KBS[nsh]; nprim; dBS, aBS [nprim each]; nao.
 */
#include <stdio.h>
#include <stdlib.h>

#include "getIntegrals.h"

static int rd_int(void) {
  int v;
  if (scanf("%d", &v) != 1) exit(3);
  return v;
}

static double rd_dbl(void) {
  double v;
  if (scanf("%lf", &v) != 1) exit(3);
  return v;
}

static int *ints(int n) {
  int *p = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) p[i] = rd_int();
  return p;
}

static double *dbls(int n) {
  double *p = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) p[i] = rd_dbl();
  return p;
}

int main(void) {
  const int nat = rd_int();
  double *geom = dbls(3 * nat);
  int *shellsECP = ints(nat);
  const int nse = rd_int();
  int *KECP = ints(nse), *lECP = ints(nse);
  const int npe = rd_int();
  double *nECP = dbls(npe), *dECP = dbls(npe), *aECP = dbls(npe);
  int *shellsBS = ints(nat);
  const int nsh = rd_int();
  int *lBS = ints(nsh), *KBS = ints(nsh);
  const int np = rd_int();
  double *dBS = dbls(np), *aBS = dbls(np);
  const int nao = rd_int();
  double *I = calloc((size_t)nao * nao, sizeof(double));
  const int rc = getIntegrals(nat, geom, shellsECP, KECP, lECP, nECP, dECP, aECP, shellsBS, lBS, KBS, dBS, aBS,
                              1024, 1e-12, 1e-14, nao, I);
  printf("%d\n", rc);
  for (int i = 0; i < nao; i++)
    for (int j = i; j < nao; j++) printf("%.17g\n", I[(size_t)i * nao + j]);
  return rc;
}
