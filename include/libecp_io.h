/* libecp_io.h -- loaders for the on-disk inputs of the reference's example
 * program, returning the flat arrays of libECP_init / getIntegrals.
 *
 * They replace the static loadGeometry / loadBS / loadECP of the
 * reference's example/ex1.c:11-123 (not part of its library), with checked
 * reads.  The ECP loader accepts both the four-column primitive records
 * ex1.c reads ("index exponent coefficient power", ex1.c:112) and the
 * three-column records of the reference's example/test_c.ecp
 * ("coefficient exponent power"), line by line.
 *
 * Each returns -1 on error (message with file and line in
 * libecp_last_error()).  Arrays are malloc'ed; release them with free().
 */
#ifndef LIBECP_IO_H
#define LIBECP_IO_H 1

#ifdef __cplusplus
extern "C" {
#endif

/* "N", a comment line, N lines "symbol x y z" (ex1.c:72-84); coordinates
   unchanged (bohr).  Returns 0. */
int libecp_load_xyz(const char *path, int *nrAtoms, double **geometry);

/* Per atom "id nShells", per shell "l K", per primitive "index exponent
   coefficient" (ex1.c:11-69).  Returns the number of shells. */
int libecp_load_basis(const char *path, int nrAtoms, int **shellsBS, int **lBS, int **KBS, double **aBS,
                      double **dBS);

/* Per atom "id nCore nShells" (0 shells: no ECP), per shell "l K", per
   primitive "index exponent coefficient power" or "coefficient exponent
   power" (ex1.c:86-123).  nCore may be NULL.  Returns the number of ECP
   shells. */
int libecp_load_ecp(const char *path, int nrAtoms, int **shellsECP, int **lECP, int **KECP, double **aECP,
                    double **dECP, double **nECP, int **nCore);

#ifdef __cplusplus
}
#endif
#endif
