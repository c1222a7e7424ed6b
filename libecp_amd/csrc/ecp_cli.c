/* ecp_integrals: the command-line counterpart of the reference's example
 * program (example/ex1.c:127-210), on the device.
 *
 *   ecp_integrals structure.xyz ECP BS            ECP matrix, shell blocks
 *   ecp_integrals structure.xyz ECP BS -e P.txt   energy  sum P V
 *   ecp_integrals structure.xyz ECP BS -g P.txt   gradient dE/dR per atom
 *
 * P.txt holds the nao x nao density as whitespace-separated numbers, row
 * after row.  The matrix is getIntegrals' (grid 1024, tolerance 1e-12,
 * accuracy 1e-14, ex1.c:160-164): its upper triangle, printed per pair of
 * shells (atom A, shell s) <= (atom B, shell t) in AO order.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/getIntegrals.h"
#include "../../include/libecp.h"
#include "../../include/libecp_ext.h"
#include "../../include/libecp_io.h"

static int ijk_dim(int l) { return (l + 1) * (l + 2) / 2; }

static double *load_matrix(const char *path, int n) {
  FILE *fp = fopen(path, "r");
  if (!fp) {
    fprintf(stderr, "ecp_integrals: cannot open %s\n", path);
    return NULL;
  }
  double *P = malloc(sizeof(double) * (size_t)n * n);
  for (size_t i = 0; i < (size_t)n * n; i++)
    if (fscanf(fp, "%lf", &P[i]) != 1) {
      fprintf(stderr, "ecp_integrals: %s: expected %d x %d numbers\n", path, n, n);
      fclose(fp);
      free(P);
      return NULL;
    }
  fclose(fp);
  return P;
}

int main(int argc, char **argv) {
  if (argc != 4 && !(argc == 6 && (!strcmp(argv[4], "-e") || !strcmp(argv[4], "-g")))) {
    fprintf(stderr, "usage: %s structure.xyz ECP BS [-e|-g density.txt]\n", argv[0]);
    return 2;
  }
  int nat = 0;
  double *geom = NULL;
  int *shECP, *lECP, *KECP, *shBS, *lBS, *KBS;
  double *aECP, *dECP, *nECP, *aBS, *dBS;
  if (libecp_load_xyz(argv[1], &nat, &geom) < 0 ||
      libecp_load_ecp(argv[2], nat, &shECP, &lECP, &KECP, &aECP, &dECP, &nECP, NULL) < 0 ||
      libecp_load_basis(argv[3], nat, &shBS, &lBS, &KBS, &aBS, &dBS) < 0) {
    fprintf(stderr, "ecp_integrals: %s\n", libecp_last_error());
    return 1;
  }
  int nsh = 0, nao = 0;
  for (int i = 0; i < nat; i++) nsh += shBS[i];
  int *aoff = malloc(sizeof(int) * (size_t)(nsh + 1));
  for (int s = 0; s < nsh; s++) {
    aoff[s] = nao;
    nao += ijk_dim(lBS[s]);
  }
  int rc = 0;
  if (argc == 6) {
    const int n = argv[4][1] == 'g';
    double *P = load_matrix(argv[5], nao);
    if (!P) return 1;
    libECPHandle *h = libECP_init(nat, geom, shECP, lECP, KECP, nECP, dECP, aECP, shBS, lBS, KBS, dBS, aBS, n, -1,
                                  NULL, 1024, 1.0E-12, 1.0E-14);
    if (!h) {
      fprintf(stderr, "ecp_integrals: %s\n", libecp_last_error());
      return 1;
    }
    double *out = calloc((size_t)3 * nat + 1, sizeof(double));
    rc = libecp_contract(h, nao, P, out);
    if (rc < 0) {
      fprintf(stderr, "ecp_integrals: %s\n", libecp_last_error());
      return 1;
    }
    if (!n) {
      printf("energy % .15e\n", out[0]);
    } else {
      for (int a = 0; a < nat; a++) printf("gradient %d % .15e % .15e % .15e\n", a, out[3 * a], out[3 * a + 1], out[3 * a + 2]);
    }
    libECP_free(h);
    free(out);
    free(P);
  } else {
    double *I = calloc((size_t)nao * nao + 1, sizeof(double));
    rc = getIntegrals(nat, geom, shECP, KECP, lECP, nECP, dECP, aECP, shBS, lBS, KBS, dBS, aBS, 1024, 1.0E-12,
                      1.0E-14, nao, I);
    if (rc) return 1;
    /* (atom, shell) of every global shell */
    int *atom = malloc(sizeof(int) * (size_t)(nsh + 1)), *local = malloc(sizeof(int) * (size_t)(nsh + 1));
    for (int i = 0, s = 0; i < nat; i++)
      for (int k = 0; k < shBS[i]; k++, s++) {
        atom[s] = i;
        local[s] = k;
      }
    printf("ECP matrix elements (upper triangle, %d AOs)\n", nao);
    for (int s = 0; s < nsh; s++)
      for (int t = s; t < nsh; t++) {
        printf("shell block (%d,%d: l=%d) x (%d,%d: l=%d)\n", atom[s], local[s], lBS[s], atom[t], local[t],
               lBS[t]);
        for (int a = 0; a < ijk_dim(lBS[s]); a++) {
          for (int b = 0; b < ijk_dim(lBS[t]); b++) {
            const int r = aoff[s] + a, c = aoff[t] + b;
            printf(" % .12e", r <= c ? I[(size_t)r * nao + c] : 0.0);
          }
          printf("\n");
        }
      }
    free(I);
    free(atom);
    free(local);
  }
  free(aoff);
  free(geom);
  free(shECP), free(lECP), free(KECP), free(aECP), free(dECP), free(nECP);
  free(shBS), free(lBS), free(KBS), free(aBS), free(dBS);
  return rc < 0 ? 1 : 0;
}
