/*
 * ref_driver.c -- driver program linked against the reference's own module
 * sources (compiled from /root/reference by oracle/build_ref.py).  It is test
 * infrastructure: it produces reference outputs for the parity tests and the
 * reference CPU timing reported by bench.py.  Lattice sizes are compile-time
 * (the reference's global.h), so one executable is built per geometry.
 *
 * Usage:
 *   ref_<tag> dump  OUTDIR [options]   write fields and operator outputs
 *   ref_<tag> bench [options]          time Dw_dble/Dw/Dwhat(_dble), print JSON
 *   ref_<tag> cnfg IN OUT [options]    import_cnfg(IN), print the plaquette
 *                                      sums, export_cnfg(OUT)
 *   ref_<tag> gcr [options]            time sap_gcr on the reference's own
 *                                      field and source, print JSON
 *   ref_<tag> dfl [options]            dfl_modes and dfl_sap_gcr on the same
 *                                      field and source, print JSON
 * Options: -bc T -cF a b -csw c -m0 m -seed s -mu mu -eoflg e
 *          -dbs b0 b1 b2 b3 -Ns n (deflation subspace), -dpro nkv nmx res
 *          nmx_gcr res_gcr (little solver), -dgen kappa mu ninv nmr ncy
 *          (subspace generation)
 *          -bs b0 b1 b2 b3 -isolv i -nmr n -ncy n -nkv n -nmx n -res r
 *          -secs t (bench: target seconds per operator)
 *          -flow n eps (dump: plaquette force of the random field to frc0,
 *          then n forward-Euler Wilson-flow steps, flowed field to ud_flow;
 *          all later outputs use the flowed field)
 *
 * The random-number sequence of "dump" is fixed:
 *   start_ranlux(0,seed); geometry(); random_ud(); [flow]; set_ud_phase();
 *   random_sd(VOLUME_TRD,2,eta,1.0) [bnd_sd2zero(ALL_PTS) for bc!=3];
 *   random_s (VOLUME_TRD,2,s32,1.0) [bnd_s2zero].
 */
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <omp.h>
#include "mpi.h"
#include "utils.h"
#include "flags.h"
#include "random.h"
#include "lattice.h"
#include "uflds.h"
#include "sflds.h"
#include "linalg.h"
#include "sw_term.h"
#include "tcharge.h"
#include "dirac.h"
#include "block.h"
#include "sap.h"
#include "mdflds.h"
#include "forces.h"
#include "wflow.h"
#include "archive.h"
#include "vflds.h"
#include "little.h"
#include "dfl.h"
#include "global.h"

static char outdir[512];

static void wr(const char *name, const void *p, size_t nbytes)
{
   char path[700];
   snprintf(path, sizeof(path), "%s/%s", outdir, name);
   FILE *f = fopen(path, "wb");
   if (f == NULL) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
   fwrite(p, 1, nbytes, f);
   fclose(f);
}

static double wtime(void) { return MPI_Wtime(); }

int main(int argc, char *argv[])
{
   int bc = 3, seed = 1, eoflg = 0, isolv = 0, nmr = 4, ncy = 5, nkv = 16, nmx = 256;
   int bs[4] = {4, 4, 4, 4}, dump = 1, i, nflow = 0;
   double eflow = 0.0;
   double cF[2] = {1.0, 1.0}, csw = 1.0, m0 = -0.5, mu = 0.0, res = 1.0e-10, secs = 2.0;
   double phi[2] = {0.0, 0.0}, phip[2] = {0.0, 0.0}, theta[3] = {0.0, 0.0, 0.0};

   if (argc < 2) { fprintf(stderr, "usage: %s dump OUTDIR|bench [options]\n", argv[0]); return 1; }
   i = 2;
   const char *cnfg_in = NULL, *cnfg_out = NULL;
   int gcr_mode = 0, dfl_mode = 0;
   int dbs[4] = {4, 4, 4, 4}, dNs = 8, dnkv = 16, dnmx = 256, dnmx_gcr = 24, dninv = 6, dnmr = 4, dncy = 5;
   double dres = 1.0e-2, dres_gcr = 1.0e-2, dkappa = 0.125, dmu = 0.0;
   if (strcmp(argv[1], "dfl") == 0)
   {
      gcr_mode = 1;
      dfl_mode = 1;
      dump = 0;
   }
   else if (strcmp(argv[1], "gcr") == 0)
   {
      /* SAP grid parameters are set below as in dump mode */
      gcr_mode = 1;
      dump = 0;
   }
   else if (strcmp(argv[1], "dump") == 0)
   {
      if (argc < 3) return 1;
      snprintf(outdir, sizeof(outdir), "%s", argv[2]);
      i = 3;
   }
   else if (strcmp(argv[1], "cnfg") == 0)
   {
      if (argc < 4) return 1;
      cnfg_in = argv[2];
      cnfg_out = argv[3];
      dump = 0;
      i = 4;
   }
   else
      dump = 0;
   for (; i < argc; i++)
   {
      if (!strcmp(argv[i], "-bc")) bc = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-cF")) { cF[0] = atof(argv[++i]); cF[1] = atof(argv[++i]); }
      else if (!strcmp(argv[i], "-csw")) csw = atof(argv[++i]);
      else if (!strcmp(argv[i], "-m0")) m0 = atof(argv[++i]);
      else if (!strcmp(argv[i], "-seed")) seed = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-mu")) mu = atof(argv[++i]);
      else if (!strcmp(argv[i], "-eoflg")) eoflg = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-bs")) { for (int k = 0; k < 4; k++) bs[k] = atoi(argv[++i]); }
      else if (!strcmp(argv[i], "-isolv")) isolv = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-nmr")) nmr = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-ncy")) ncy = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-nkv")) nkv = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-nmx")) nmx = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-res")) res = atof(argv[++i]);
      else if (!strcmp(argv[i], "-secs")) secs = atof(argv[++i]);
      else if (!strcmp(argv[i], "-flow")) { nflow = atoi(argv[++i]); eflow = atof(argv[++i]); }
      else if (!strcmp(argv[i], "-dbs")) { for (int k = 0; k < 4; k++) dbs[k] = atoi(argv[++i]); }
      else if (!strcmp(argv[i], "-Ns")) dNs = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-dpro"))
      {
         dnkv = atoi(argv[++i]); dnmx = atoi(argv[++i]); dres = atof(argv[++i]);
         dnmx_gcr = atoi(argv[++i]); dres_gcr = atof(argv[++i]);
      }
      else if (!strcmp(argv[i], "-dgen"))
      {
         dkappa = atof(argv[++i]); dmu = atof(argv[++i]); dninv = atoi(argv[++i]);
         dnmr = atoi(argv[++i]); dncy = atoi(argv[++i]);
      }
   }

   mpi_init(argc, argv);
   set_lat_parms(6.0, 1.0, 0, NULL, 0, csw);
   set_bc_parms(bc, 1.0, 1.0, cF[0], cF[1], phi, phip, theta);
   start_ranlux(0, seed);
   geometry();
   set_sw_parms(m0);
   set_tm_parms(eoflg);
   if ((dump || gcr_mode) && (((L0 / bs[0]) * (L1 / bs[1]) * (L2 / bs[2]) * (L3 / bs[3])) % (2 * NTHREAD)) == 0)
      set_sap_parms(bs, isolv, nmr, ncy);

   alloc_wsd(10);
   alloc_ws(2 * nkv + 9 + (dfl_mode ? dNs + 4 : 0));
   if (dfl_mode)
   {
      set_dfl_parms(dbs, dNs);
      set_dfl_pro_parms(dnkv, dnmx, dres, dnmx_gcr, dres_gcr);
      set_dfl_gen_parms(dkappa, dmu, dninv, dnmr, dncy);
      alloc_wv(2 * dnmx_gcr + 3);
      alloc_wvd(2 * dnkv + 6);
   }
   spinor_dble **wsd = reserve_wsd(8);
   spinor **ws = reserve_ws(8);

   if (cnfg_in)
   {
      /* ---------------- configuration file round trip -------------------- */
      import_cnfg((char *)cnfg_in, 0);
      printf("{\"plaq_sum\": %.17e, \"plaq_wsum\": %.17e}\n", plaq_sum_dble(1), plaq_wsum_dble(1));
      export_cnfg((char *)cnfg_out);
      MPI_Finalize();
      return 0;
   }

   random_ud();
   if (gcr_mode)
   {
      /* ---------------- SAP+GCR timing (devel/sap/check3.c style) ------- */
      double t1, t2, rho, rq;
      int ifail = 0, status = 0;
      qflt q0, q1;
      if (nflow > 0) fwd_euler(nflow, eflow);
      set_ud_phase();
      if ((((L0 / bs[0]) * (L1 / bs[1]) * (L2 / bs[2]) * (L3 / bs[3])) % (2 * NTHREAD)) != 0)
      {
         printf("{\"error\": \"SAP block grid not allowed for NTHREAD=%d\"}\n", NTHREAD);
         MPI_Finalize();
         return 1;
      }
      random_sd(VOLUME_TRD, 2, wsd[0], 1.0);
      bnd_sd2zero(ALL_PTS, wsd[0]);
      if (dfl_mode)
      {
         int dif[2] = {0, 0}, dst[6] = {0, 0, 0, 0, 0, 0};
         double t0 = wtime();
         dfl_modes(dif, dst);
         t1 = wtime();
         rho = dfl_sap_gcr(nkv, nmx, 0, res, mu, wsd[0], wsd[1], dif + 1, dst + 2);
         t2 = wtime();
         Dw_dble(mu, wsd[1], wsd[2]);
         mulr_spinor_add_dble(VOLUME_TRD, 2, wsd[2], wsd[0], -1.0);
         q1 = norm_square_dble(VOLUME_TRD, 3, wsd[2]);
         q0 = norm_square_dble(VOLUME_TRD, 3, wsd[0]);
         rq = sqrt(q1.q[0] / q0.q[0]);
         printf("{\"nthread\": %d, \"volume\": %d, \"modes_ifail\": %d, \"modes_status\": [%d, %d], "
                "\"dfl_ifail\": %d, \"dfl_status\": [%d, %d, %d], \"dfl_rho\": %.6e, "
                "\"true_residual\": %.6e, \"modes_sec\": %.6e, \"dfl_sec\": %.6e}\n",
                NTHREAD, VOLUME, dif[0], dst[0], dst[1], dif[1], dst[2], dst[3], dst[4], rho, rq, t1 - t0, t2 - t1);
         MPI_Finalize();
         return 0;
      }
      t1 = wtime();
      rho = sap_gcr(nkv, nmx, 0, res, mu, wsd[0], wsd[1], &ifail, &status);
      t2 = wtime();
      Dw_dble(mu, wsd[1], wsd[2]);
      mulr_spinor_add_dble(VOLUME_TRD, 2, wsd[2], wsd[0], -1.0);
      q1 = norm_square_dble(VOLUME_TRD, 3, wsd[2]);
      q0 = norm_square_dble(VOLUME_TRD, 3, wsd[0]);
      rq = sqrt(q1.q[0] / q0.q[0]);
      printf("{\"nthread\": %d, \"volume\": %d, \"gcr_status\": %d, \"gcr_ifail\": %d, "
             "\"gcr_rho\": %.6e, \"true_residual\": %.6e, \"gcr_sec\": %.6e}\n",
             NTHREAD, VOLUME, status, ifail, rho, rq, t2 - t1);
      MPI_Finalize();
      return 0;
   }
   if (dump && nflow > 0)
   {
      plaq_frc();
      wr("frc0.f64", (*mdflds()).frc, 4 * VOLUME * sizeof(su3_alg_dble));
      fwd_euler(nflow, eflow);
      wr("ud_flow.f64", udfld(), 4 * VOLUME * sizeof(su3_dble));
   }
   if (dump)
   {
      /* the field before the phase: plaquette sums and the exported file */
      double pq[2] = {plaq_sum_dble(1), plaq_wsum_dble(1)};
      char path[700];
      wr("plaq.f64", pq, sizeof(pq));
      snprintf(path, sizeof(path), "%s/cnfg.dat", outdir);
      export_cnfg(path);
   }
   set_ud_phase();

   if (!dump)
   {
      /* ---------------- timing of the operators (OpenMP, NTHREAD) ------ */
      double wt, t1, t2;
      int nc, k;
      sw_term(NO_PTS);
      assign_ud2u();
      assign_swd2sw();
      random_sd(VOLUME_TRD, 2, wsd[0], 1.0);
      bnd_sd2zero(ALL_PTS, wsd[0]);
      assign_sd2s(VOLUME_TRD, 2, wsd[0], ws[0]);
      printf("{\"nthread\": %d, \"volume\": %d", NTHREAD, VOLUME);
      for (k = 0; k < 4; k++)
      {
         const char *nm[4] = {"Dw_dble", "Dw", "Dwhat_dble", "Dwhat"};
         if (k == 2) { sw_term(ODD_PTS); assign_swd2sw(); }
         nc = 0;
         wt = 0.0;
         t1 = wtime();
         while (wt < secs)
         {
            if (k == 0) Dw_dble(mu, wsd[0], wsd[1]);
            else if (k == 1) Dw((float)mu, ws[0], ws[1]);
            else if (k == 2) Dwhat_dble(mu, wsd[0], wsd[1]);
            else Dwhat((float)mu, ws[0], ws[1]);
            nc++;
            t2 = wtime();
            wt = t2 - t1;
         }
         printf(", \"%s_calls\": %d, \"%s_sec\": %.6e", nm[k], nc, nm[k], wt);
      }
      printf("}\n");
      MPI_Finalize();
      return 0;
   }

   /* ------------------------------ dump --------------------------------- */
   int V = VOLUME;
   wr("ipt.i32", ipt, sizeof(int) * V);
   wr("ud.f64", udfld(), sizeof(su3_dble) * 4 * V);

   spinor_dble *eta = wsd[0], *sd1 = wsd[1], *sd2 = wsd[2], *sd3 = wsd[3];
   random_sd(VOLUME_TRD, 2, eta, 1.0);
   bnd_sd2zero(ALL_PTS, eta);
   wr("eta.f64", eta, sizeof(spinor_dble) * V);
   {
      char path[700];
      snprintf(path, sizeof(path), "%s/eta_sfld0.dat", outdir);
      export_sfld(path, 0, eta);
      snprintf(path, sizeof(path), "%s/eta_sfld1.dat", outdir);
      export_sfld(path, 1, eta);
   }
   random_s(VOLUME_TRD, 2, ws[0], 1.0f);
   bnd_s2zero(ALL_PTS, ws[0]);
   wr("s32.f32", ws[0], sizeof(spinor) * V);

   u3_alg_dble **ft = ftensor();
   for (int n = 0; n < 6; n++)
   {
      char nm[32];
      snprintf(nm, sizeof(nm), "ft%d.f64", n);
      wr(nm, ft[n], sizeof(u3_alg_dble) * V);
   }

   sw_term(NO_PTS);
   wr("swd.f64", swdfld(), sizeof(pauli_dble) * 2 * V);
   assign_ud2u();
   assign_swd2sw();

   /* FP64 operator pieces; inputs may be modified (written back as *_in) */
   struct { const char *name; int kind; } ops[5] = {
      {"Dw_dble", 0}, {"Dwee_dble", 1}, {"Dwoo_dble", 2}, {"Dwoe_dble", 3}, {"Dweo_dble", 4}};
   for (int k = 0; k < 5; k++)
   {
      char nm[64];
      assign_sd2sd(VOLUME_TRD, 2, eta, sd1);
      random_sd(VOLUME_TRD, 2, sd2, 1.0);    /* pre-fill of the output */
      snprintf(nm, sizeof(nm), "%s_r0.f64", ops[k].name);
      wr(nm, sd2, sizeof(spinor_dble) * V);
      if (ops[k].kind == 0) Dw_dble(mu, sd1, sd2);
      else if (ops[k].kind == 1) Dwee_dble(mu, sd1, sd2);
      else if (ops[k].kind == 2) Dwoo_dble(mu, sd1, sd2);
      else if (ops[k].kind == 3) Dwoe_dble(sd1, sd2);
      else Dweo_dble(sd1, sd2);
      snprintf(nm, sizeof(nm), "%s_in.f64", ops[k].name);
      wr(nm, sd1, sizeof(spinor_dble) * V);
      snprintf(nm, sizeof(nm), "%s.f64", ops[k].name);
      wr(nm, sd2, sizeof(spinor_dble) * V);
   }

   /* FP32 operator */
   assign_s2s(VOLUME_TRD, 2, ws[0], ws[1]);
   Dw((float)mu, ws[1], ws[2]);
   wr("Dw_in.f32", ws[1], sizeof(spinor) * V);
   wr("Dw.f32", ws[2], sizeof(spinor) * V);

   /* Schur complement (odd part inverted) */
   int ifail_inv = sw_term(ODD_PTS);
   wr("swd_oinv.f64", swdfld(), sizeof(pauli_dble) * 2 * V);
   assign_sd2sd(VOLUME_TRD, 2, eta, sd1);
   set_sd2zero(VOLUME_TRD, 2, sd2);
   Dwhat_dble(mu, sd1, sd2);
   wr("Dwhat_dble.f64", sd2, sizeof(spinor_dble) * V);
   assign_swd2sw();
   assign_s2s(VOLUME_TRD, 2, ws[0], ws[1]);
   set_s2zero(VOLUME_TRD, 2, ws[2]);
   Dwhat((float)mu, ws[1], ws[2]);
   wr("Dwhat.f32", ws[2], sizeof(spinor) * V);

   /* linear algebra: double-double dot products */
   complex_qflt cq = spinor_prod_dble(VOLUME_TRD, 3, eta, wsd[1]);
   qflt nq = norm_square_dble(VOLUME_TRD, 3, eta);
   double la[6] = {cq.re.q[0], cq.re.q[1], cq.im.q[0], cq.im.q[1], nq.q[0], nq.q[1]};
   wr("qsum.f64", la, sizeof(la));
   complex z32 = spinor_prod(VOLUME_TRD, 3, ws[0], ws[2]);
   float n32 = norm_square(VOLUME_TRD, 3, ws[0]);
   float lf[3] = {z32.re, z32.im, n32};
   wr("salg.f32", lf, sizeof(lf));

   /* the SAP block grid needs an even multiple of NTHREAD blocks
      (flags/sap_parms.c:105); otherwise the SAP parts are left out */
   int nblk = (L0 / bs[0]) * (L1 / bs[1]) * (L2 / bs[2]) * (L3 / bs[3]);
   int ifail_bgr = 0, ifail = 0, status = 0, dosap = (nblk % (2 * NTHREAD)) == 0;
   double rho = 0.0, t1 = 0.0, t2 = 0.0;
   qflt q1 = {{0.0, 0.0}}, q0 = {{1.0, 0.0}};
   if (!dosap) goto info;

   /* SAP sweep: ncy cycles on (psi=0, rho=s32) */
   sw_term(NO_PTS);
   assign_ud2u();
   assign_swd2sw();
   alloc_bgr(SAP_BLOCKS);
   assign_ud2ubgr(SAP_BLOCKS);
   ifail_bgr = assign_swd2swbgr(SAP_BLOCKS, isolv ? ODD_PTS : NO_PTS);
   set_s2zero(VOLUME_TRD, 2, ws[3]);
   assign_s2s(VOLUME_TRD, 2, ws[0], ws[4]);
   for (int n = 0; n < ncy; n++) sap((float)mu, isolv, nmr, ws[3], ws[4]);
   wr("sap_psi.f32", ws[3], sizeof(spinor) * V);
   wr("sap_rho.f32", ws[4], sizeof(spinor) * V);

   /* block MinRes on block 0: s[1] = rho restricted to the block */
   {
      int nb, isw;
      block_t *b = blk_list(SAP_BLOCKS, &nb, &isw);
      if (isolv == 0)
      {
         assign_s2sblk(SAP_BLOCKS, 0, ALL_PTS, ws[0], 1);
         wr("blk0_s1_in.f32", b[0].s[1], sizeof(spinor) * b[0].vol);
         Dw_blk(SAP_BLOCKS, 0, (float)mu, 1, 2);
         wr("blk0_Dw.f32", b[0].s[2], sizeof(spinor) * b[0].vol);
         blk_mres(0, (float)mu, nmr);
         wr("blk0_s0.f32", b[0].s[0], sizeof(spinor) * b[0].vol);
         wr("blk0_s1.f32", b[0].s[1], sizeof(spinor) * b[0].vol);
      }
      wr("blk0_imb.i32", b[0].imb, sizeof(int) * b[0].vol);
   }

   /* SAP+GCR */
   t1 = wtime();
   rho = sap_gcr(nkv, nmx, 0, res, mu, eta, sd3, &ifail, &status);
   t2 = wtime();
   wr("gcr_psi.f64", sd3, sizeof(spinor_dble) * V);
   Dw_dble(mu, sd3, sd1);
   mulr_spinor_add_dble(VOLUME_TRD, 2, sd1, eta, -1.0);
   q1 = norm_square_dble(VOLUME_TRD, 3, sd1);
   q0 = norm_square_dble(VOLUME_TRD, 3, eta);
info:
   {
      char path[700];
      snprintf(path, sizeof(path), "%s/info.json", outdir);
      FILE *f = fopen(path, "w");
      fprintf(f, "{\"V\": %d, \"NTHREAD\": %d, \"sap\": %d, \"ifail_inv\": %d, \"ifail_bgr\": %d, "
                 "\"gcr_ifail\": %d, \"gcr_status\": %d, \"gcr_rho\": %.17e, "
                 "\"gcr_check\": %.17e, \"gcr_sec\": %.6e}\n",
              V, NTHREAD, dosap, ifail_inv, ifail_bgr, ifail, status, rho, sqrt(q1.q[0] / q0.q[0]),
              t2 - t1);
      fclose(f);
   }
   MPI_Finalize();
   return 0;
}
