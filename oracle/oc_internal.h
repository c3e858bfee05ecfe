/*
 * oc_internal.h -- shared state of the CPU oracle (test infrastructure only).
 */
#ifndef OC_INTERNAL_H
#define OC_INTERNAL_H

#include <stdint.h>
#include "oracle.h"

typedef struct
{
   int64_t st[13][4];  /* 12 words x 4 lanes of 48-bit state + carry row */
   int pr, ir;
} oc_rlx;

typedef struct
{
   oc_rlx s;
   int is, is_old;
   double rd[48];
   float rs[96];
   int dbl;            /* 1: ranlxd stream, 0: ranlxs stream */
} oc_stream;

typedef struct
{
   int bo[4], vol;
   int *ipt;           /* block lex -> block index          */
   int (*iup)[4], (*idn)[4];  /* block index neighbours (vol = outside) */
   int *imb;           /* block index -> global index       */
} oc_block;

struct oc_ctx
{
   int L[4], Ltrd[4];
   int V, VT, nthread;
   int *ipt, *lex;     /* lex -> index, index -> lex */
   int (*iup)[4], (*idn)[4];
   int *x0;            /* index -> time coordinate */

   int bc;
   double cF[2];
   double m0, csw;
   int eoflg;
   int sap_bs[4], sap_isolv, sap_nmr, sap_ncy;

   oc_stream *rs, *rd;   /* per-thread ranlxs / ranlxd streams */
   int rlx_init;

   double *ud;   /* 4V su3_dble in reference layout */
   float *u;     /* FP32 copy */
   double *swd;  /* 2V pauli_dble */
   float *sw;    /* FP32 copy */
   int swd_inv[2];  /* even/odd parts hold inverses */
   int ud_phase;

   /* SAP block grid */
   int nb, nbh, bvol;
   oc_block *blk;
   float *sw_bgr;   /* FP32 Pauli for block solver, odd part inverted if isolv */
   double *swd_bgr; /* FP64 version of it (odd part inverted, not rounded) */
   int sap_ready;
};

/* indices of neighbours in the reference link layout */
static inline int oc_link_fwd(const oc_ctx *c, int ix, int mu)
{
   /* U(x,mu): for odd x it is ud[8*(ix-V/2)+2mu]; for even x it is the
      backward link of the odd site x+mu (README.global, uflds.c) */
   int h = c->V / 2;
   if (ix >= h) return 8 * (ix - h) + 2 * mu;
   return 8 * (c->iup[ix][mu] - h) + 2 * mu + 1;
}

static inline int oc_link_bwd(const oc_ctx *c, int ix, int mu)
{
   /* U(x-mu,mu) */
   int h = c->V / 2;
   if (ix >= h) return 8 * (ix - h) + 2 * mu + 1;
   return 8 * (c->idn[ix][mu] - h) + 2 * mu;
}

void oc_gauss_internal(oc_ctx *c, int thread, double *r, int n);
void oc_build_blocks(oc_ctx *c);

#endif
