/* getIntegrals.h -- one-call API of the MI355X ECP-integral library.
 *
 * Drop-in for the reference's getIntegrals.h:7-13 (definition
 * getIntegrals.c:45-95).  Note the argument order here: (shellsECP, KECP,
 * lECP, ...), unlike libECP_init.  Adds the full ECP matrix into the upper
 * triangle (row <= column) of the caller's row-major matrix I with row stride
 * rowdim; rows/columns are the Cartesian AOs, atom-major, shell by shell,
 * libint order inside a shell.  I must be zeroed by the caller; it is
 * accumulated into (+=).  Returns 1 if the handle cannot be created, else 0
 * (as the reference, also when an integration failed; use the handle API to
 * see that status).
 */
#ifndef GET_INTEGRALS_H
#define GET_INTEGRALS_H 1

#ifdef __cplusplus
extern "C" {
#endif

int getIntegrals(int nrAtoms, double *geometry,
                 int *shellsECP, int *KECP,
                 int *lECP, double *nECP, double *dECP, double *aECP,
                 int *shellsBS, int *lBS, int *KBS,
                 double *dBS, double *aBS,
                 int largeGridOrder, double tolerance, double accuracy,
                 int rowdim, double *I);

#ifdef __cplusplus
}
#endif
#endif
