/* Loaders for the on-disk inputs of the reference's example program
 * (example/ex1.c:11-123): an .xyz geometry, a basis-set file and an ECP
 * file, returned as the flat arrays libECP_init / getIntegrals take.
 *
 * Formats (whitespace-separated numbers, one record per line; blank lines
 * are skipped):
 *   .xyz  N / a comment line / N lines "symbol x y z" (ex1.c:72-84; the
 *         coordinates are passed on unchanged, i.e. in bohr)
 *   basis per atom "id nShells", per shell "l K", per primitive
 *         "index exponent coefficient" (ex1.c:11-69)
 *   ECP   per atom "id nCore nShells" (nShells = 0: no ECP), per shell
 *         "l K", per primitive either "index exponent coefficient power"
 *         (the four columns ex1.c:112 reads) or "coefficient exponent
 *         power" (the three columns of the reference's example/test_c.ecp,
 *         which ex1.c's four-column read misparses); each line is read by
 *         its own column count.
 * Unlike ex1.c these loaders check every read and report the file and line
 * of the first malformed record.  The arrays are malloc'ed; free() them.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/libecp_io.h"

void ecp_api_set_error(const char *msg);

typedef struct {
  FILE *fp;
  const char *path;
  int line;
  char buf[1024];
} reader;

static int fail(reader *r, const char *what) {
  char msg[1200];
  snprintf(msg, sizeof(msg), "%s:%d: %s", r->path, r->line, what);
  ecp_api_set_error(msg);
  return -1;
}

/* next non-blank line into r->buf; 0, or -1 at end of file */
static int next_line(reader *r) {
  while (fgets(r->buf, sizeof(r->buf), r->fp)) {
    r->line++;
    const char *p = r->buf;
    while (*p && isspace((unsigned char)*p)) p++;
    if (*p) return 0;
  }
  return -1;
}

/* up to max numbers of the next non-blank line; the count, or -1 at end of
   file, or -2 when a field is not a number */
static int numbers(reader *r, double *v, int max) {
  if (next_line(r)) return -1;
  int n = 0;
  char *p = r->buf, *end;
  for (;;) {
    while (*p && isspace((unsigned char)*p)) p++;
    if (!*p) break;
    if (n == max) return -2;
    errno = 0;
    v[n] = strtod(p, &end);
    if (end == p || errno == ERANGE) return -2;
    n++;
    p = end;
  }
  return n;
}

static int is_int(double x, int lo, int hi) { return x == (double)(int)x && x >= lo && x <= hi; }

static int open_reader(reader *r, const char *path) {
  memset(r, 0, sizeof(*r));
  r->path = path;
  r->fp = fopen(path, "r");
  if (!r->fp) {
    char msg[1200];
    snprintf(msg, sizeof(msg), "%s: cannot open (%s)", path, strerror(errno));
    ecp_api_set_error(msg);
    return -1;
  }
  return 0;
}

int libecp_load_xyz(const char *path, int *nrAtoms, double **geometry) {
  reader r;
  double v[4];
  *geometry = NULL;
  if (open_reader(&r, path)) return -1;
  int n = numbers(&r, v, 1);
  if (n != 1 || !is_int(v[0], 1, 1 << 24)) {
    fclose(r.fp);
    return fail(&r, "expected the atom count");
  }
  const int nat = (int)v[0];
  /* the comment line may be empty */
  if (!fgets(r.buf, sizeof(r.buf), r.fp)) {
    fclose(r.fp);
    return fail(&r, "missing comment line");
  }
  r.line++;
  double *g = malloc(sizeof(double) * 3 * (size_t)nat);
  if (!g) {
    fclose(r.fp);
    return fail(&r, "out of memory");
  }
  for (int i = 0; i < nat; i++) {
    if (next_line(&r)) {
      free(g);
      fclose(r.fp);
      return fail(&r, "fewer atoms than announced");
    }
    char *p = r.buf, *end;
    while (*p && isspace((unsigned char)*p)) p++;
    while (*p && !isspace((unsigned char)*p)) p++; /* element symbol */
    for (int k = 0; k < 3; k++) {
      errno = 0;
      g[3 * i + k] = strtod(p, &end);
      if (end == p || errno == ERANGE) {
        free(g);
        fclose(r.fp);
        return fail(&r, "expected 'symbol x y z'");
      }
      p = end;
    }
  }
  fclose(r.fp);
  *nrAtoms = nat;
  *geometry = g;
  return 0;
}

/* growable arrays; a failed allocation sets oom and drops the value */
typedef struct {
  int *v;
  size_t n, cap;
  int oom;
} ivec;
typedef struct {
  double *v;
  size_t n, cap;
  int oom;
} dvec;

static void ipush(ivec *a, int x) {
  if (a->n == a->cap) {
    const size_t cap = a->cap ? 2 * a->cap : 16;
    int *v = realloc(a->v, sizeof(int) * cap);
    if (!v) {
      a->oom = 1;
      return;
    }
    a->v = v;
    a->cap = cap;
  }
  a->v[a->n++] = x;
}

static void dpush(dvec *a, double x) {
  if (a->n == a->cap) {
    const size_t cap = a->cap ? 2 * a->cap : 16;
    double *v = realloc(a->v, sizeof(double) * cap);
    if (!v) {
      a->oom = 1;
      return;
    }
    a->v = v;
    a->cap = cap;
  }
  a->v[a->n++] = x;
}

/* "l K" of a shell */
static int shell_header(reader *r, int *l, int *K) {
  double v[2];
  const int n = numbers(r, v, 2);
  if (n != 2 || !is_int(v[0], 0, 64) || !is_int(v[1], 1, 1 << 20)) return fail(r, "expected 'l K' of a shell");
  *l = (int)v[0];
  *K = (int)v[1];
  return 0;
}

int libecp_load_basis(const char *path, int nrAtoms, int **shellsBS, int **lBS, int **KBS, double **aBS,
                      double **dBS) {
  reader r;
  ivec sh = {0}, l = {0}, K = {0};
  dvec a = {0}, d = {0};
  int rc = -1;
  if (open_reader(&r, path)) return -1;
  for (int i = 0; i < nrAtoms; i++) {
    double v[3];
    int n = numbers(&r, v, 2);
    if (n != 2 || !is_int(v[1], 0, 1 << 20)) {
      fail(&r, n == -1 ? "fewer atoms than the geometry" : "expected 'id nShells' of an atom");
      goto out;
    }
    const int ns = (int)v[1];
    ipush(&sh, ns);
    for (int s = 0; s < ns; s++) {
      int ls, ks;
      if (shell_header(&r, &ls, &ks)) goto out;
      ipush(&l, ls);
      ipush(&K, ks);
      for (int p = 0; p < ks; p++) {
        if (numbers(&r, v, 3) != 3) {
          fail(&r, "expected 'index exponent coefficient'");
          goto out;
        }
        dpush(&a, v[1]);
        dpush(&d, v[2]);
      }
    }
  }
  rc = (int)l.n;
  if (sh.oom || l.oom || K.oom || a.oom || d.oom) rc = fail(&r, "out of memory");
out:
  fclose(r.fp);
  if (rc < 0) {
    free(sh.v), free(l.v), free(K.v), free(a.v), free(d.v);
    return -1;
  }
  *shellsBS = sh.v ? sh.v : calloc(1, sizeof(int));
  *lBS = l.v ? l.v : calloc(1, sizeof(int));
  *KBS = K.v ? K.v : calloc(1, sizeof(int));
  *aBS = a.v ? a.v : calloc(1, sizeof(double));
  *dBS = d.v ? d.v : calloc(1, sizeof(double));
  return rc;
}

int libecp_load_ecp(const char *path, int nrAtoms, int **shellsECP, int **lECP, int **KECP, double **aECP,
                    double **dECP, double **nECP, int **nCore) {
  reader r;
  ivec sh = {0}, l = {0}, K = {0}, core = {0};
  dvec a = {0}, d = {0}, pw = {0};
  int rc = -1;
  if (open_reader(&r, path)) return -1;
  for (int i = 0; i < nrAtoms; i++) {
    double v[4];
    int n = numbers(&r, v, 3);
    if (n != 3 || !is_int(v[1], 0, 1 << 20) || !is_int(v[2], 0, 64)) {
      fail(&r, n == -1 ? "fewer atoms than the geometry" : "expected 'id nCore nShells' of an atom");
      goto out;
    }
    ipush(&core, (int)v[1]);
    const int ns = (int)v[2];
    ipush(&sh, ns);
    for (int s = 0; s < ns; s++) {
      int ls, ks;
      if (shell_header(&r, &ls, &ks)) goto out;
      ipush(&l, ls);
      ipush(&K, ks);
      for (int p = 0; p < ks; p++) {
        n = numbers(&r, v, 4);
        if (n == 4) { /* index exponent coefficient power (ex1.c:112) */
          dpush(&a, v[1]);
          dpush(&d, v[2]);
          dpush(&pw, v[3]);
        } else if (n == 3) { /* coefficient exponent power (example/test_c.ecp) */
          dpush(&d, v[0]);
          dpush(&a, v[1]);
          dpush(&pw, v[2]);
        } else {
          fail(&r, "expected 'index exponent coefficient power' or 'coefficient exponent power'");
          goto out;
        }
      }
    }
  }
  rc = (int)l.n;
  if (sh.oom || l.oom || K.oom || core.oom || a.oom || d.oom || pw.oom) rc = fail(&r, "out of memory");
out:
  fclose(r.fp);
  if (rc < 0) {
    free(sh.v), free(l.v), free(K.v), free(core.v), free(a.v), free(d.v), free(pw.v);
    return -1;
  }
  *shellsECP = sh.v ? sh.v : calloc(1, sizeof(int));
  *lECP = l.v ? l.v : calloc(1, sizeof(int));
  *KECP = K.v ? K.v : calloc(1, sizeof(int));
  *aECP = a.v ? a.v : calloc(1, sizeof(double));
  *dECP = d.v ? d.v : calloc(1, sizeof(double));
  *nECP = pw.v ? pw.v : calloc(1, sizeof(double));
  if (nCore)
    *nCore = core.v ? core.v : calloc(1, sizeof(int));
  else
    free(core.v);
  return rc;
}
