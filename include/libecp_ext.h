/* libecp_ext.h -- additive entry points of the MI355X ECP-integral library.
 *
 * None of these has a counterpart in the reference (chrr/libECP); they sit
 * beside the drop-in API of libecp.h / getIntegrals.h and never change it.
 */
#ifndef LIBECP_EXT_H
#define LIBECP_EXT_H 1

#include "libecp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host-side screening of a handle's molecule (no device work): returns the
   number of (C, shell a, shell b >= a) triplets that calculateECPIntegrals
   evaluates, in its callback order; when keys != NULL, writes up to cap
   records of 5 ints (C, A, sa, B, sb).  Returns -1 on error. */
int libecp_screened_triplets(libECPHandle *h, int *keys, int cap);

/* N_ECP * N_sh (N_sh + 1) / 2: the reference's triplet loop count before
   screening (libecp.c:256-312), the unit of the throughput metric. */
long long libecp_nominal_triplets(libECPHandle *h);

#ifdef __cplusplus
}
#endif
#endif
