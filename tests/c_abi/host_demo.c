/*
 * host_demo.c -- a C99 host program written against include/openqcd_hip.h
 * only, the way a reference main program (e.g. devel/sap/check3.c) calls the
 * solver: geometry and parameters, the reference's random field, the SW
 * term, workspace fields, Dw_dble and SAP+GCR with the recomputed residual.
 * Built and linked against libopenqcd_hip.so by tests/test_c_abi.py (CPU:
 * compile, link, run "symbols"; GPU: run "solve").
 *
 *   host_demo symbols   print the number of entry points taken from the
 *                       library (no GPU call)
 *   host_demo solve     8^4 periodic lattice, csw=1.0, m0=-0.5, seed 1:
 *                       prints status, ifail and the true residual
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "openqcd_hip.h"

#include "host_demo_symbols.h"  /* generated: the table of entry points */

int main(int argc, char *argv[])
{
   if (argc < 2 || strcmp(argv[1], "solve") != 0)
   {
      int n = 0, k;
      for (k = 0; k < (int)(sizeof(oqh_symbols) / sizeof(oqh_symbols[0])); k++)
         n += (oqh_symbols[k] != NULL);
      printf("symbols %d\n", n);
      return 0;
   }

   int L[4] = {8, 8, 8, 8}, bs[4] = {4, 4, 4, 4}, ifail = 0, status = 0;
   const int vol = 8 * 8 * 8 * 8;
   if (oqh_init(0) || oqh_geometry(L, L)) return 1;
   oqh_set_bc_parms(3, 1.0, 1.0);
   oqh_set_sw_parms(-0.5, 1.0);
   oqh_set_tm_parms(0);
   oqh_set_sap_parms(bs, 1, 4, 5);
   start_ranlux(0, 1);
   random_ud();
   set_ud_phase();
   if (sw_term(NO_PTS)) return 2;

   alloc_wsd(3);
   spinor_dble **wsd = reserve_wsd(3);
   random_sd(vol, 0, wsd[0], 1.0);
   double rho = sap_gcr(16, 256, 0, 1e-10, 0.0, wsd[0], wsd[1], &ifail, &status);
   Dw_dble(0.0, wsd[1], wsd[2]);
   mulr_spinor_add_dble(vol, 0, wsd[2], wsd[0], -1.0);
   qflt r2 = norm_square_dble(vol, 0, wsd[2]);
   qflt e2 = norm_square_dble(vol, 0, wsd[0]);
   double res = sqrt((r2.q[0] + r2.q[1]) / (e2.q[0] + e2.q[1]));
   printf("status %d ifail %d rho %.3e true_residual %.3e\n", status, ifail, rho, res);
   release_wsd();
   oqh_finalize();
   return (ifail == 0 && res <= 1.2e-10) ? 0 : 3;
}
