// driver.hip -- quark propagators in the manner of the reference's main
// program ms4 (main/ms4.c).
//
// For one gauge field (the current one, e.g. after import_cnfg and
// set_ud_phase as ms4.c:606-607 do) nsrc random sources on time slice x0 are
// drawn as random_source (ms4.c:383-406) does, solved with SAP+GCR and the
// current SAP parameters (solve_dirac, SAP_GCR branch, ms4.c:409-459) and
// written in the universal quark-field format (save_prop, export_sfld
// branch, ms4.c:470-481) as <dir>/<nbase>n<icnfg>.s<isrc>.  The sources
// do not depend on the solutions (the solver draws no random numbers), so
// they are drawn first and solved together, up to 16 per sap_gcr_multi call.
#include <cstdio>
#include <vector>

#include "oqh_internal.h"

using namespace oqh;

extern "C" int oqh_ms4_propagators(int x0, int nsrc, int nkv, int nmx, int istop, double res, double mu,
                                   const char *dir, const char *nbase, int icnfg, int *status)
{
   const char *where = "oqh_ms4_propagators [ms4.c]";
   Lib &L = lib();
   if (!require_geometry(where)) return -1;
   if (nsrc < 1 || x0 < 0 || x0 >= L.g.L[0] || dir == nullptr || nbase == nullptr || status == nullptr)
   {
      fail(where, "Parameters are out of range");
      return -1;
   }
   if (L.bc == 0 && (x0 == 0 || x0 == L.g.L[0] - 1))
   {
      fail(where, "the source time slice is on the boundary of open boundary conditions");
      return -1;
   }
   constexpr int NB = 16;  // sources per sap_gcr_multi call
   const int nb = nsrc < NB ? nsrc : NB;
   alloc_wsd(L.wsd_used + 2 * nb);
   spinor_dble **w = reserve_wsd(2 * nb);
   if (!w) return -1;
   int nfail = 0;
   std::vector<double> rho(nb);
   std::vector<int> ifail(nb);
   char path[1024];
   for (int i0 = 0; i0 < nsrc; i0 += nb)
   {
      const int n = (nsrc - i0) < nb ? (nsrc - i0) : nb;
      for (int k = 0; k < n; k++) oqh_random_source(x0, w[k]);
      sap_gcr_multi(n, nkv, nmx, istop, res, mu, w, w + nb, rho.data(), ifail.data(), status + i0);
      for (int k = 0; k < n; k++)
      {
         if (ifail[k] < 0)
         {
            nfail++;
            char msg[128];
            snprintf(msg, sizeof(msg), "SAP_GCR solver failed (source %d, ifail %d)", i0 + k, ifail[k]);
            fail(where, msg);
            continue;
         }
         snprintf(path, sizeof(path), "%s/%sn%d.s%d", dir, nbase, icnfg, i0 + k);
         export_sfld(path, 0, w[nb + k]);
      }
   }
   release_wsd();
   return nfail;
}
