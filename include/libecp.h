/* libecp.h -- handle API of the MI355X ECP-integral library.
 *
 * Drop-in for the reference's public header (chrr/libECP src/libecp.h:1-31):
 * same type names, same function names and signatures, same return codes.
 * Two corrections to the reference header's documentation, not to its ABI:
 *  - libECP_init's arguments 4 and 5 are (lECP, KECP), the order of the
 *    reference's definition (libecp.c:54-61) and of its only caller
 *    (getIntegrals.c:78); the reference header names them the other way round
 *    (libecp.h:16-17).  Positional meaning is what binds.
 *  - the callback is invoked as cb(A, sa, la, shifta, B, sb, lb, shiftb, C,
 *    block, args) (libecp.c:372); the reference typedef's parameter names
 *    (libecp.h:8-12) do not match that order.
 */
#ifndef LIBECP_H
#define LIBECP_H 1

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _libECPHandle libECPHandle;

/* Receives one block per (C, shell a, shell b >= a, shift, type): an
   IJK(la+shifta) x IJK(lb+shiftb) row-major array (for n = 2 the (1,0) and
   (0,1) shifts keep the unshifted l: IJK(la) x IJK(lb)) of Cartesian integrals in
   libint order (p: x,y,z; d: xx,xy,xz,yy,yz,zz; ...).  The library owns the
   block; it is valid only during the call.  Calls arrive on the caller's
   thread in the reference order C -> A -> B >= A -> sa -> sb -> shift ->
   type 1 (local) then type 2 (semi-local). */
typedef void (*ECPCallback)(int A, int sa, int la, int shifta,
                            int B, int sb, int lb, int shiftb,
                            int C, double *I, void *p);

/* Replaces libecp.h:15-21 (definition libecp.c:54-61).  Geometry in bohr,
   atom-major xyz.  Basis: shellsBS[atom], then per shell lBS/KBS, then per
   primitive dBS/aBS (used unnormalised).  ECP: shellsECP[atom] (0 = none),
   per ECP shell lECP/KECP, per primitive nECP (radial power including the
   r^2 Jacobian), dECP, aECP; the highest l of an atom is its local channel.
   n: derivative order 0, 1 or 2 (basis l + n at most 6); for n > 0 each
   triplet is delivered once per angular-momentum shift of the reference's
   table (libecp.c:243-250), with the shifts in the callback.
   lmax/shellOrdering: custom Cartesian ordering (NULL = libint).  Returns
   NULL on invalid input or when the Bessel tabulation does not converge. */
libECPHandle *libECP_init(int nrAtoms, double *geometry,
                          int *shellsECP, int *lECP, int *KECP,
                          double *nECP, double *dECP, double *aECP,
                          int *shellsBS, int *lBS, int *KBS,
                          double *dBS, double *aBS,
                          int n, int lmax, int *shellOrdering,
                          int largeGridOrder, double tolerance, double accuracy);

/* Replaces libecp.h:23-27 (definition libecp.c:213-404).
   Returns 0 on success, 1 if a type-1 integration failed, 2 if a type-2
   integration failed; blocks before the failing one (in the order above)
   have been delivered.  Returns -1 on a device error. */
int calculateECPIntegrals(libECPHandle *h, ECPCallback cb, void *args);

/* Replaces libecp.h:29 (definition libecp.c:407-444). */
void libECP_free(libECPHandle *h);

/* Additive helper with no counterpart in the reference: the text of the last
   error raised by this library on the calling thread ("" if none). */
const char *libecp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
