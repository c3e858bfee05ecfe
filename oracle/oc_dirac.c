/*
 * oc_dirac.c -- oracle: Dw, Dw_dble and their even/odd pieces.
 * TEST INFRASTRUCTURE ONLY (see oracle.h and oc_dirac_tmpl.h).
 */
#include <stdlib.h>
#include <string.h>
#include "oc_internal.h"

#define R double
#define SFX _dble
#define UFLD(c) ((c)->ud)
#define SWFLD(c) ((c)->swd)
#define MULP oc_mul_pauli_dble
#include "oc_dirac_tmpl.h"

/* double hopping term shared with the FP64 block operators in oc_sap.c */
void oc_hop_term_d(int mu, int sgn, int dag, const double *u, const double *v, double coef, double *r)
{
   hop_term_dble(mu, sgn, dag, u, v, coef, r);
}
#undef R
#undef SFX
#undef UFLD
#undef SWFLD
#undef MULP

#define R float
#define SFX
#define UFLD(c) ((c)->u)
#define SWFLD(c) ((c)->sw)
#define MULP oc_mul_pauli
#include "oc_dirac_tmpl.h"
#undef R
#undef SFX
#undef UFLD
#undef SWFLD
#undef MULP

/* float hopping term shared with the block operators in oc_sap.c */
void oc_hop_term_f(int mu, int sgn, int dag, const float *u, const float *v, float coef, float *r)
{
   hop_term(mu, sgn, dag, u, v, coef, r);
}
